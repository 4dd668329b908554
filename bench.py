#!/usr/bin/env python3
"""Benchmark: attempted spin updates/s of the Metropolis sweep with parallel tempering on one MI355X.

Default workload (``--config batch_c3``): 256 independent PT chains advanced together --
lattice SC / BCC (2 basis) / FCC (4 basis) and L in {8, 12, 16, 24, 32, 48} drawn per chain,
16-32 replicas on a geometric 0.5-2.5 J ladder, J1 = 1, J2 ~ U[-0.3, 0.3], 'random' proposals,
uniform-on-sphere initial spins, all from master seed 7 (472,136,384 spin updates per sweep of
the batch).  One step = one sweep of every replica of every chain + one measurement of every
replica + one exchange round of every chain.  Synthetic inputs, resident in HBM before the
timed region.

Other workloads: pt_fcc16 (FCC 16^3, J1-J3 + Delta, 64 replicas), mc_sc8, pt_sc64_dm (SC 64^3,
J + B + DM + Delta, 32 replicas), pt_fcc32 (FCC 32^3, J1-J2 + B + Delta, 64 replicas),
learned_fcc32 (learned proposal vs plain Metropolis, effective samples/s of E), copy_probe.

Prints one JSON line (stdout).  Extra legs:
  roofline      -- the dominant sweep kernel: minimal algorithmic bytes of its launches (each
                   spin a colour phase reads, read once; each updated spin, written once) / the
                   launches' summed duration, timed with HIP events on the engine's stream;
                   traffic = HBM bytes per launch from the committed rocprofv3 counter summary
  cpu_baseline  -- the oracle's reference-order sweep (MT19937, sequential sites, FP64),
                   one process per host core, on a bounded sample of the same workload
  extras        -- (batch_c3) short runs of pt_fcc16 / pt_sc64_dm / pt_fcc32, the copy
                   bandwidth, and the learned-proposal ESS ratios of the committed profile
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "attempted spin updates/s (all replicas); effective samples/s of E vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from espins_amd import lattice as Lt
    if name == "pt_fcc16":
        m, tems = Lt.config_c2(L=16, R=64)
        desc = ("PT FCC 16^3 cells (N=16384), J1=-1 J2=+0.2 J3=+0.05, Delta=-0.1 (z), "
                "64 replicas geometric T 0.05-1.5 J, exchange every sweep")
        return m, tems, desc, 4242
    if name == "mc_sc8":
        m, tems = Lt.config_c1(L=8)
        return m, tems, "plain MC SC 8^3 (N=512), J1=1, 16 T lin 0.5-2.0", 2024
    if name == "pt_sc64_dm":
        m, tems = Lt.config_c4(L=64, R=32)
        return m, tems, "PT SC 64^3 (N=262144) J1=1 + DM 0.3 + B 0.05 + Delta=+0.02, 32 replicas", 77
    if name == "pt_fcc32":
        m, tems = Lt.config_c5(L=32, R=64)
        return m, tems, "PT FCC 32^3 cells (N=131072) J1=-1 J2=0.1 B=0.1 Delta=-0.05, 64 replicas", 5
    raise SystemExit(f"unknown config {name}")


def bytes_per_update(m):
    """Algorithmic HBM bytes per attempted update, CSR/ELL gather model:
    52 + sum over distinct bonds of (28 + 8 p_b), p_b = FP64 coupling scalars on the bond
    (J 1, B 1, D 3); own spin read+write 48 B + row offset 4 B; per bond neighbour spin 24 B +
    int32 index 4 B."""
    import numpy as np
    N = m.N
    tot = 0.0
    for i in range(0, N, max(1, N // 512)):     # sample of sites (periodic lattices)
        bonds = {}
        for tab, p in ((m.J, 1), (m.B, 1), (m.D, 3)):
            if tab is None:
                continue
            for l in range(int(tab.cnt[i])):
                j = int(tab.nbr[l, i])
                bonds[j] = bonds.get(j, 0) + p
        tot += 52 + sum(28 + 8 * p for p in bonds.values())
    return tot / len(range(0, N, max(1, N // 512)))


def phase_bytes(m):
    """Algorithmic HBM bytes of the colour phases of one sweep of one replica, minimal-traffic
    model: a colour phase reads every spin it needs once (its own sites and all their bond
    partners, 24 B each) and writes every site it updates once (24 B).  Couplings, geometry
    rows and per-basis constants are per-lattice tables shared by every replica (count 0).
    Returns (bytes per sweep, updates per sweep = N, number of colours); the colouring is the
    library's own (plan_model)."""
    import numpy as np
    from espins_amd.engine import plan_model
    ncol, col, _ = plan_model(m)
    col = np.asarray(col)
    parts = []
    for tab in (m.J, m.B, m.D):
        if tab is None or tab.z == 0:
            continue
        nb = np.asarray(tab.nbr) - 1                          # [z][N], 0-based, -1 = empty
        cnt = np.asarray(tab.cnt)
        ok = np.arange(tab.z)[:, None] < cnt[None, :]
        parts.append((nb, ok))
    total = 0
    for c in range(ncol):
        act = col == c
        need = act.copy()
        for nb, ok in parts:
            sel = ok & act[None, :]
            need[nb[sel]] = True
        total += 24 * int(need.sum()) + 24 * int(act.sum())
    return total, m.N, ncol


# ---------------------------------------------------------------- CPU baseline ------------
def _cpu_worker(args):
    name, T, seconds, seed = args
    import numpy as np
    from espins_amd import lattice as Lt
    from oracle import oracle as O
    if name == "pt_fcc16":
        m = Lt.cubic_model("fcc", 16, J=(-1.0, 0.2, 0.05), K=-0.1, builder=O.build_table_fn)
    elif name == "mc_sc8":
        m = Lt.cubic_model("sc", 8, J=(1.0,), builder=O.build_table_fn)
    elif name == "pt_sc64_dm":
        m = Lt.cubic_model("sc", 64, J=(1.0,), B=0.05, D=0.3, K=0.02, builder=O.build_table_fn)
    else:
        m = Lt.cubic_model("fcc", 32, J=(-1.0, 0.1), B=0.1, K=-0.05, builder=O.build_table_fn)
    o = O.Oracle(m)
    s = Lt.random_spins(m.N, 1, seed)[0]
    st = O.mt_state(seed)
    e = o.energy(s)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        e, _ = o.sweep_ref(s, 1.0 / T, 0, 0.125, st, 1, e)
        n += 1
    return n * m.N, time.perf_counter() - t0


def cpu_baseline(name, tems, seconds):
    import multiprocessing as mp
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = max(1, min(ncores, 16))
    jobs = [(name, float(tems[k % len(tems)]), seconds, 1000 + k) for k in range(ncores)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(ncores) as pool:
        res = pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t0
    upd = sum(r[0] for r in res)
    busy = max(r[1] for r in res)
    return dict(value=upd / busy, unit="updates/s", cores=ncores, kind="port",
                sample=(f"oracle reference-order Metropolis (MT19937, sites 1..N, FP64) on the "
                        f"same Hamiltonian, one process per core, one temperature each, "
                        f"{seconds:.0f} s per core ({upd} updates, {wall:.1f} s wall)"))


# ---------------------------------------------------------------- batched workload --------
def _cpu_worker_c3(args):
    kind, L, J2, T, seconds, seed = args
    from espins_amd import lattice as Lt
    from oracle import oracle as O
    m = Lt.cubic_model(kind, L, J=(1.0, J2), builder=O.build_table_fn)
    o = O.Oracle(m)
    s = Lt.random_spins(m.N, 1, seed)[0]
    st = O.mt_state(seed)
    e = o.energy(s)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        e, _ = o.sweep_ref(s, 1.0 / T, 0, 0.125, st, 1, e)
        n += 1
    return (kind, L), n * m.N, time.perf_counter() - t0


def cpu_baseline_c3(chains, seconds):
    """The oracle's reference-order sweep on the batch's own update mix: every lattice class
    (kind, L) of the batch is timed on host cores (one process per core, all cores busy at
    once, a chain of that class at a mid-ladder temperature), and the batch's CPU throughput is
    the update-weighted harmonic mean of the class rates times the cores -- the rate at which
    the host cores would complete one sweep of every replica of every chain."""
    import multiprocessing as mp
    ncores = max(1, min(len(os.sched_getaffinity(0)), 16))
    share, first = {}, {}
    for i, (kind, L, J2, R, tems) in enumerate(chains):
        nb = {"sc": 1, "bcc": 2, "fcc": 4}[kind]
        share[(kind, L)] = share.get((kind, L), 0) + nb * L ** 3 * R
        first.setdefault((kind, L), (kind, L, J2, float(tems[len(tems) // 2]), seconds, 100 + i))
    classes = sorted(share, key=lambda c: -share[c])
    jobs = [first[c] for c in classes]
    while len(jobs) % ncores:          # keep every core busy in every round
        jobs.append(first[classes[len(jobs) % len(classes)]])
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(ncores) as pool:
        res = pool.map(_cpu_worker_c3, jobs)
    wall = time.perf_counter() - t0
    rate = {}
    for c, u, t in res:
        rate.setdefault(c, []).append(u / t)
    rate = {c: sum(v) / len(v) for c, v in rate.items()}
    total = sum(share.values())
    per_core = total / sum(share[c] / rate[c] for c in classes)
    return dict(value=per_core * ncores, unit="updates/s", cores=ncores, kind="port",
                per_core=per_core,
                sample=(f"oracle reference-order Metropolis (MT19937, sites 1..N, FP64) on every one of "
                        f"the batch's {len(classes)} lattice classes (one process per core, {ncores} cores "
                        f"busy at once, {seconds:.1f} s per class), combined as the update-weighted "
                        f"harmonic mean of the class rates x cores ({wall:.1f} s wall)"),
                class_rates={f"{k}{L}": r for (k, L), r in rate.items()})


def run_batch(args):
    """256 independent PT chains (SC/BCC/FCC, L in {8..48}, 16-32 replicas, J1 = 1, random J2)
    advanced together; one step = one sweep + measurement + exchange of every chain."""
    import multiprocessing as mp
    from espins_amd import lattice as Lt
    chains = Lt.config_c3(n_chains=args.chains, seed=7)
    cpu = None
    if not args.no_cpu:    # before any GPU call (forked workers must not inherit a GPU context)
        log(f"cpu_baseline: every lattice class of the batch, {args.cpu_seconds / 4:.1f} s each ...")
        cpu = cpu_baseline_c3(chains, args.cpu_seconds / 4)
        log(f"cpu_baseline: {cpu['value'] / 1e6:.2f} M updates/s on {cpu['cores']} cores "
            f"({cpu['per_core'] / 1e6:.2f} M/s per core)")
    from espins_amd.engine import Batch
    t0 = time.perf_counter()
    b = Batch(0)
    total = 0
    for i, (kind, L, J2, R, tems) in enumerate(chains):
        m = Lt.cubic_model_fast(kind, L, J=(1.0, J2))
        b.add_chain(m, tems, Lt.random_spins(m.N, R, 1000 + i))
        total += m.N * R
    b.commit()
    log(f"batch: {len(chains)} chains, {total} spin updates per sweep, set up in "
        f"{time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    if args.batch_path == "auto":
        b.autotune(nsweeps=4, cache=args.tune_cache)
    else:                     # every chain on one kernel path, one lane per site
        for k in range(len(chains)):
            b.set_chain_config(k, 1 if args.batch_path == "global" else 0, 1)
    shapes = {}
    for k in range(len(chains)):
        i = b.info(k)
        key = (i["lds"], i["tps"], i["chunk"])
        shapes[key] = shapes.get(key, 0) + 1
    log(f"autotune: {time.perf_counter() - t1:.1f} s, (lds, lanes, chunk) -> chains {shapes}")
    graph = not args.no_graph
    b.run(args.warmup, measure=True, exchange=True, seed=7, graph=graph)
    b.reset_stats()
    b.sync()
    t2 = time.perf_counter()
    b.run(args.steps, measure=True, exchange=True, seed=7, sweep0=args.warmup, step0=args.warmup,
          graph=graph)
    b.sync()
    dt = time.perf_counter() - t2
    value = total * args.steps / dt
    rej = sum(int(b.stats(k)["rejected"].sum()) for k in range(len(chains)))
    log(f"timed: {args.steps} steps in {dt:.3f} s -> {value / 1e9:.3f} G updates/s, "
        f"acceptance {1 - rej / (total * args.steps):.4f}")

    # roofline of the dominant kernel (the global-memory colour phases, k_bphase): minimal
    # algorithmic bytes of the timed launches / their summed duration (HIP events)
    t3 = time.perf_counter()
    cls = {}
    alg_sweep = gat_sweep = 0.0
    for k, (kind, L, J2, R, tems) in enumerate(chains):
        if b.info(k)["lds"]:
            continue
        if (kind, L) not in cls:
            m = Lt.cubic_model_fast(kind, L, J=(1.0, J2))
            cls[(kind, L)] = (phase_bytes(m)[0], bytes_per_update(m), m.N)
        pb, gb, N = cls[(kind, L)]
        alg_sweep += pb * R
        gat_sweep += gb * N * R
    log(f"phase-byte model: {len(cls)} lattice classes, {time.perf_counter() - t3:.1f} s")
    roof = None
    if alg_sweep > 0:
        nsw = 3
        tm = b.time_sweeps(nsw, mode="random", seed=11, sweep0=10 ** 9)
        sec = tm["kernel_ms"] * 1e-3
        achieved = alg_sweep * nsw / sec / 1e9
        upd_s = tm["updates"] / sec
        log(f"k_bphase: {tm['nlaunch']} launches, mean {tm['kernel_ms'] / tm['nlaunch'] * 1e3:.1f} us, "
            f"{alg_sweep / (tm['updates'] / nsw):.1f} B/update algorithmic -> {achieved:.0f} GB/s; "
            f"kernel-only {upd_s / 1e9:.3f} G updates/s")
        traffic = measured_traffic("batch_c3", "k_bphase")
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic,
                "traffic_source": ("profiles/pmc_k_bphase_batch_c3.txt: FETCH_SIZE + WRITE_SIZE "
                                   "per dispatch, copy-kernel calibrated, mean over the k_bphase "
                                   "dispatches of the same workload") if traffic else None,
                "kernel": "k_bphase",
                "algorithmic_bytes_per_launch": alg_sweep * nsw / tm["nlaunch"],
                "bytes_per_update": alg_sweep * nsw / tm["updates"],
                "model": ("minimal colour-phase traffic: each spin a phase reads once, each "
                          "updated spin written once, 24 B per FP64 3-vector"),
                "updates_per_launch": tm["updates"] / tm["nlaunch"],
                "mean_launch_us": tm["kernel_ms"] / tm["nlaunch"] * 1e3,
                "kernel_updates_per_s": upd_s,
                "gather_model": {"bytes_per_update": gat_sweep * nsw / tm["updates"],
                                 "GBs": gat_sweep * nsw / sec / 1e9}}
    out = {
        "metric": METRIC, "value": value, "unit": "updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"{len(chains)} independent PT chains, SC/BCC/FCC, L in "
                                "{8,12,16,24,32,48}, 16-32 replicas on geometric 0.5-2.5 J ladders, "
                                "J1=1, J2~U[-0.3,0.3], exchange every sweep"),
                   "chains": len(chains), "updates_per_step": total, "graph": graph,
                   "batch_path": args.batch_path,
                   "kernel_shapes": {f"lds={k[0]},lanes={k[1]}": v for k, v in shapes.items()}},
        "roofline": roof, "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    b.close()
    if not args.no_extras:
        extras_into(out, args)
    print(json.dumps(out), flush=True)


def profile_ess():
    """Learned-proposal ESS ratios from the committed profile of the learned_fcc32 run."""
    import ast
    path = os.path.join(ROOT, "profiles", "ess_learned_vs_plain_fcc32.txt")
    if not os.path.exists(path):
        return None
    got = {}
    for line in open(path):
        t = line.strip()
        for key in ("ess_ratio_vs_random", "ess_ratio_vs_const", "ess_ratio_M_vs_random", "ess_ratio_M_vs_const"):
            if t.startswith(key + " ="):
                got[key] = ast.literal_eval(t.split("=", 1)[1].strip())
    if not got:
        return None
    got["source"] = "profiles/ess_learned_vs_plain_fcc32.txt (python bench.py --config learned_fcc32)"
    return got


def extras_into(out, args):
    """Short single-ladder runs of the other workloads and the copy bandwidth, as extra keys
    of the batched line (no CPU legs; the same sweep path and timing as --config <name>)."""
    from espins_amd.engine import Engine
    ex = {}
    for cfg, steps, warm in (("pt_fcc16", 2000, 200), ("pt_sc64_dm", 200, 20), ("pt_fcc32", 300, 30)):
        r = run_single(args, cfg, steps, warm, with_cpu=False, emit=False)
        rf = r["roofline"]
        ex[cfg] = {"workload": r["config"]["workload"], "value": r["value"], "unit": "updates/s",
                   "ms_per_step": r["ms_per_step"], "steps": steps, "kernel": rf["kernel"],
                   "mean_launch_us": rf["mean_launch_us"], "roofline_frac": rf["frac"],
                   "bytes_per_update": rf["bytes_per_update"]}
    out["extras"] = ex
    ess = profile_ess()
    if ess:
        out["learned_ess_ratio"] = ess
    eng = Engine(0)
    copy_gbs = 2 * (1 << 28) / (eng.copy_probe(1 << 28, reps=10) * 1e-3) / 1e9
    eng.close()
    if out.get("roofline"):
        out["roofline"]["copy_bandwidth_GBs"] = copy_gbs
        out["roofline"]["frac_of_copy"] = out["roofline"]["achieved"] / copy_gbs
    log(f"copy bandwidth {copy_gbs:.0f} GB/s")


# ---------------------------------------------------------------- learned proposal --------
def run_learned(args):
    """Frustrated FCC 32^3 PT (N = 131072, 64 replicas): train the learned proposal on local
    conditionals of equilibrated plain chains, then compare effective samples/s of E of the
    learned chain with plain Metropolis ('random' and 'const' proposals) on the same schedule
    (one sweep + measurement + exchange per step) from the same start."""
    import numpy as np
    from espins_amd import lattice as Lt
    from espins_amd import learn as Ln
    m, tems, desc, seed = workload("pt_fcc32")
    cpu = None
    if not args.no_cpu:
        log(f"cpu_baseline: {args.cpu_seconds:.0f} s per core ...")
        cpu = cpu_baseline("pt_fcc32", tems, args.cpu_seconds)
        log(f"cpu_baseline: {cpu['value'] / 1e6:.2f} M updates/s on {cpu['cores']} cores")
    from espins_amd.engine import Engine
    R = len(tems)
    eng = Engine(0)
    eng.set_model(m)
    # Ordered start (four-sublattice tetrahedral state, E/N = -2.25 J, the lowest of the
    # simple orders of this Hamiltonian): the traces then measure equilibrium autocorrelation
    # at the cold end instead of domain coarsening from a random start.
    init = (Lt.sublattice_spins(m, Lt.FCC_TETRAHEDRAL, R) if args.start == "ordered"
            else Lt.random_spins(m.N, R, seed, m.mag))
    eng.set_replicas(tems, init)
    graph = not args.no_graph

    def pt(n, mode, sweep0, tilt=0.0):
        eng.run(n, steps_per_round=1, steps_measure=1, measure=True, exchange=True,
                start_parity=1 + (sweep0 % 2), mode=mode, tilt=tilt, seed=seed, sweep0=sweep0,
                step0=sweep0, graph=graph)

    t0 = time.perf_counter()
    pt(args.warmup, "random", 0)
    eng.sync()
    log(f"equilibration: {args.warmup} PT steps, {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    losses = Ln.train_proposal(eng, n_records=args.records, collect_every=2,
                               per_collect=max(1, args.records // 16), epochs=args.epochs,
                               batch=4096, lr=2e-3, seed=seed, mode="random", sweep0=args.warmup)
    t_train = time.perf_counter() - t1
    log(f"training: {args.records} records, {args.epochs} epochs, loss {losses[0]:.4f} -> "
        f"{losses[-1]:.4f}, {t_train:.1f} s (collection sweeps included)")
    start = eng.spins(by_slot=True)
    sweep0 = args.warmup + 64
    res = {}
    for mode, tilt in (("random", 0.0), ("const", args.tilt), ("learned", 0.0)):
        eng.set_replicas(tems, start)
        pt(args.ess_warmup, mode, sweep0, tilt)
        eng.reset_stats()
        eng.set_trace(args.steps)
        eng.sync()
        t2 = time.perf_counter()
        pt(args.steps, mode, sweep0 + args.ess_warmup, tilt)
        eng.sync()
        dt = time.perf_counter() - t2
        trace = eng.trace()
        tr = trace[:, :, 0]                       # E per measurement and slot
        st = eng.stats()
        eng.set_trace(0)
        ess, tau = Ln.ess_per_slot(tr)
        ess_m, tau_m = Ln.ess_per_slot(trace[:, :, 1])   # |m|
        acc = 1.0 - st["rejected"] / float(m.N * args.steps)
        h = args.steps // 2       # drift: first-half minus second-half mean, in standard errors
        se = np.sqrt(2 * tau / h) * tr.std(0)
        drift = (tr[:h].mean(0) - tr[h:].mean(0)) / np.maximum(se * np.sqrt(2), 1e-300)
        res[mode] = dict(ess_s=ess / dt, tau=tau, dt=dt, acc=acc, e=tr.mean(0), drift=drift,
                         ess_m_s=ess_m / dt, tau_m=tau_m)
        log(f"{mode:8s}: {args.steps} steps in {dt:.2f} s ({m.N * R * args.steps / dt / 1e9:.3f} G updates/s), "
            f"acceptance {acc.min():.4f}-{acc.max():.4f}, tau_int(E) median {np.median(tau):.1f} "
            f"max {tau.max():.1f} sweeps, ESS/s total {ess.sum() / dt:.1f}, "
            f"tau_int(|m|) median {np.median(tau_m):.1f}, ESS/s of |m| total {ess_m.sum() / dt:.1f}, "
            f"half-to-half drift max {np.abs(drift).max():.1f} sigma")
        log("   tau by slot (cold -> hot): " + " ".join(f"{x:.0f}" for x in tau[::4]))
    lo = slice(0, R // 4)                       # coldest quarter of the ladder
    ratio = {k: res["learned"]["ess_s"] / res[k]["ess_s"] for k in ("random", "const")}
    ratio_m = {k: res["learned"]["ess_m_s"] / res[k]["ess_m_s"] for k in ("random", "const")}
    reliable = int(np.sum(res["random"]["tau"] * 50 < args.steps))
    tm = eng.time_sweeps(20, mode="learned", seed=seed + 1, sweep0=10 ** 9)
    B = phase_bytes(m)[0] / m.N                  # minimal bytes per update, mean over colours
    ms_launch = tm["kernel_ms"] / tm["nlaunch"]
    achieved = tm["updates_per_launch"] * B / (ms_launch * 1e-3) / 1e9
    log(f"k_phase_learned: mean {ms_launch * 1e3:.1f} us per launch, "
        f"{tm['updates_per_launch']} updates/launch -> {achieved:.0f} GB/s algorithmic")
    L = res["learned"]
    out = {
        "metric": METRIC, "value": float(L["ess_s"].sum()), "unit": "effective samples of E/s (sum over slots)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": L["dt"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": desc + ", learned vMF-mixture proposal vs plain Metropolis", "N": m.N,
                   "replicas": R, "records": args.records, "epochs": args.epochs, "graph": graph,
                   "const_tilt": args.tilt, "ess_warmup": args.ess_warmup, "start": args.start},
        "ess_ratio_vs_random": {"median": float(np.median(ratio["random"])),
                                "coldest_quarter": float(np.median(ratio["random"][lo])),
                                "min": float(ratio["random"].min())},
        "ess_ratio_vs_const": {"median": float(np.median(ratio["const"])),
                               "coldest_quarter": float(np.median(ratio["const"][lo])),
                               "min": float(ratio["const"].min())},
        "ess_ratio_M_vs_random": {"median": float(np.median(ratio_m["random"])),
                                  "coldest_quarter": float(np.median(ratio_m["random"][lo])),
                                  "min": float(ratio_m["random"].min())},
        "ess_ratio_M_vs_const": {"median": float(np.median(ratio_m["const"])),
                                 "coldest_quarter": float(np.median(ratio_m["const"][lo])),
                                 "min": float(ratio_m["const"].min())},
        "tau_int_E": {k: [float(x) for x in v["tau"]] for k, v in res.items()},
        "tau_int_M": {k: [float(x) for x in v["tau_m"]] for k, v in res.items()},
        "random_tau_reliable_slots": reliable,
        "drift_sigma_max": {k: float(np.abs(v["drift"]).max()) for k, v in res.items()},
        "updates_per_s": {k: m.N * R * args.steps / v["dt"] for k, v in res.items()},
        "acceptance_mean": {k: float(v["acc"].mean()) for k, v in res.items()},
        "train": {"loss_first": float(losses[0]), "loss_last": float(losses[-1]), "seconds": t_train},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_phase_learned",
                     "bytes_per_update": B, "model": "minimal colour-phase traffic",
                     "updates_per_launch": tm["updates_per_launch"],
                     "mean_launch_us": ms_launch * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    eng.close()


# ---------------------------------------------------------------- counter calibration ------
def run_copy_probe(args):
    """Plain device copy of a known size (k_copy): the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters into HBM bytes."""
    from espins_amd.engine import Engine
    n = 1 << 28
    eng = Engine(0)
    ms = eng.copy_probe(n, reps=max(1, args.steps))
    gbs = 2 * n / (ms * 1e-3) / 1e9
    log(f"k_copy: {n} B read + {n} B written per launch, {ms * 1e3:.1f} us -> {gbs:.0f} GB/s")
    print(json.dumps({"metric": "device copy bandwidth (read + write)", "value": gbs, "unit": "GB/s",
                      "n_gpus": 1, "steps": args.steps, "bytes_read_per_launch": n,
                      "bytes_written_per_launch": n, "ms_per_launch": ms}), flush=True)
    eng.close()


def measured_traffic(config, kernel):
    """HBM bytes per launch of `kernel` on `config`, from the committed rocprofv3 --pmc summary
    profiles/pmc_<kernel>_<config>.txt (line 'traffic_bytes_per_launch = X'), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{config}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


# ---------------------------------------------------------------- main ---------------------
STEP_DEFAULTS = {"batch_c3": (200, 20), "learned_fcc32": (2000, 2000), "copy_probe": (3, 0)}


def enable_observables(eng, m, config, which="all"):
    """The optional observables of the reference with their default sizes: order parameter
    'op' (axes along z), staggered moment 'sm' (alternating signs over the basis), 'pm' P(m),
    'sc' spin correlations over 2 shells, and 'sf' S(q) on the default 51 x 51 mesh spanned by
    b1 and b3 every 100 measurements (parameters.F90:440-592)."""
    sel = {"op", "sm", "pm", "sc", "sf"} if which == "all" else set(which.split(","))
    import numpy as np
    from espins_amd import lattice as Lt
    kind = next(k for k in ("fcc", "bcc", "sc") if k in config)
    L = int("".join(ch for ch in config.split(kind)[1] if ch.isdigit()) or 0)
    nb = m.nb
    if "op" in sel:
        eng.set_orderparam(np.tile([0.0, 0.0, 1.0], (nb, 1)))
    if "sm" in sel:
        eng.set_staggered(np.array([1.0 if a % 2 == 0 else -1.0 for a in range(nb)]))
    if "pm" in sel:
        eng.set_pm(True)
    if "sc" in sel:
        cnt, nbr = Lt.shell_tables(kind, L, 2)
        eng.set_spincorr(np.zeros(m.N, np.int32), cnt, nbr)
    if "sf" in sel:
        b = 2.0 * np.pi * np.eye(3)                # reciprocal vectors of the cubic cell
        ij = np.array([(i, j) for i in range(51) for j in range(51)], float) / 50.0
        q = ij[:, :1] * b[0] + ij[:, 1:] * b[2]
        eng.set_sfactor(q, Lt.site_positions(kind, L), np.array([1.0, 0.0, 0.0]), every=100)
    log(f"observables on: {sorted(sel)} (S(q): 2601 q-points every 100 measurements)")


def run_single(args, config, steps, warmup, with_cpu=True, emit=True):
    """One single-context workload (plain MC or PT of one replica ladder); returns the JSON
    line as a dict (and prints it when emit)."""
    m, tems, desc, seed = workload(config)
    cpu = None
    if with_cpu:   # before any GPU call (forked workers must not inherit a GPU context)
        log(f"cpu_baseline: {args.cpu_seconds:.0f} s per core ...")
        cpu = cpu_baseline(config, tems, args.cpu_seconds)
        log(f"cpu_baseline: {cpu['value'] / 1e6:.2f} M updates/s on {cpu['cores']} cores")

    from espins_amd import lattice as Lt
    from espins_amd.engine import Engine
    # No torch here: its wheel carries its own HIP runtime, and a second runtime in the
    # process keeps libespins_hip.so from opening the device.  The barrier is the engine's
    # stream synchronisation (every kernel of the path runs on that one stream).

    R = len(tems)
    eng = Engine(0)
    eng.set_model(m)
    eng.set_path(args.path)
    eng.set_tps(args.tps)
    eng.set_replicas(tems, Lt.random_spins(m.N, R, seed, m.mag))
    if getattr(args, "observables", None):
        enable_observables(eng, m, config, args.observables)
    info = eng.model_info()
    log(f"{config} model: {info}")
    pt = R > 1 and config.startswith("pt")
    graph = not args.no_graph

    def run(n, sweep0, step0):
        if pt:    # one round = one sweep + measurement + exchange (pt_steps_swap = 1)
            eng.run(n, steps_per_round=1, steps_measure=1, measure=True, exchange=True,
                    start_parity=1 + (step0 % 2), mode="random", seed=seed, sweep0=sweep0,
                    step0=step0, graph=graph)
        else:     # independent temperatures: all n measured steps in one round
            eng.run(1, steps_per_round=n, steps_measure=1, measure=True, exchange=False,
                    mode="random", seed=seed, sweep0=sweep0, graph=False)

    run(warmup, 0, 0)
    eng.reset_stats()
    eng.sync()
    t0 = time.perf_counter()
    run(steps, warmup, warmup)
    eng.sync()
    dt = time.perf_counter() - t0
    updates = m.N * R * steps
    value = updates / dt
    st = eng.stats()
    acc_rate = 1.0 - st["rejected"].sum() / float(updates)
    sr, ns = eng.pt_map()
    swap_rate = float(ns[:, 1].sum()) / max(1, ns[:, 0].sum())
    log(f"{config} timed: {steps} steps in {dt:.3f} s -> {value / 1e9:.3f} G updates/s; "
        f"acceptance {acc_rate:.4f}, swap acceptance {swap_rate:.4f}")

    # sweep-kernel roofline (HIP events around each launch on the engine's stream)
    nsw = max(10, min(200, steps // 10))
    tm = eng.time_sweeps(nsw, mode="random", seed=seed + 1, sweep0=10 ** 9)
    pb = phase_bytes(m)[0] * R                  # minimal bytes of one sweep of every replica
    ms_launch = tm["kernel_ms"] / tm["nlaunch"]
    launches_per_sweep = tm["nlaunch"] / nsw
    alg_launch = pb / launches_per_sweep
    achieved = alg_launch / (ms_launch * 1e-3) / 1e9
    kernel = "k_lds" if info["lds"] else "k_phase"
    G = bytes_per_update(m)
    log(f"{kernel}: {tm['nlaunch']} launches, mean {ms_launch * 1e3:.2f} us, "
        f"{tm['updates_per_launch']} updates/launch, {pb / (m.N * R):.0f} B/update minimal "
        f"-> {achieved:.0f} GB/s")
    zJ, zB, zD = m.J.z, (m.B.z if m.B else 0), (m.D.z if m.D else 0)
    flops = 70 + 6 * zJ + 12 * zB + 11 * zD
    traffic = measured_traffic(config, kernel)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": dt / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "N": m.N, "replicas": R, "sweeps_per_step": 1,
                   "exchange_every": 1 if pt else 0, "mode": "random", "colours": info["ncol"],
                   "kernel_path": "lds" if info["lds"] else "global", "graph": graph,
                   "observables": getattr(args, "observables", None)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": (f"profiles/pmc_{kernel}_{config}.txt (FETCH_SIZE + WRITE_SIZE, "
                                        "copy-kernel calibrated)") if traffic else None,
                     "kernel": kernel, "bytes_per_update": pb / (m.N * R),
                     "model": "minimal colour-phase traffic (each spin a phase reads once, each "
                              "updated spin written once, 24 B per FP64 3-vector)",
                     "updates_per_launch": tm["updates_per_launch"],
                     "mean_launch_us": ms_launch * 1e3,
                     "gather_model": {"bytes_per_update": G,
                                      "GBs": tm["updates_per_launch"] * G / (ms_launch * 1e-3) / 1e9},
                     "fp64_flops_per_update": flops, "fp64_frac": value * flops / 78.6e12},
        "cpu_baseline": cpu,
        "acceptance": acc_rate,
        "swap_acceptance": swap_rate,
    }
    if emit:
        print(json.dumps(out), flush=True)
    eng.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default per config)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default per config)")
    ap.add_argument("--config", default="batch_c3",
                    choices=("batch_c3", "pt_fcc16", "mc_sc8", "pt_sc64_dm", "pt_fcc32",
                             "learned_fcc32", "copy_probe"))
    ap.add_argument("--path", type=int, default=0, help="0 auto, 1 global-memory phases, 2 LDS")
    ap.add_argument("--tps", type=int, default=0, help="lanes per site (0 automatic)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--observables", default=None,
                    help="single-ladder configs: optional observables measured on the device, 'all' or a "
                         "comma list of op, sm, pm, sc, sf")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="batch_c3: skip the short single-ladder runs reported as extra keys")
    ap.add_argument("--chains", type=int, default=256, help="batch_c3: number of chains")
    ap.add_argument("--tune-cache", default=None, help="batch_c3: autotuner cache file")
    ap.add_argument("--batch-path", default="global", choices=("auto", "global", "fit"),
                    help="batch_c3: every chain on the global-memory path (default), autotuned "
                         "kernel shapes, or LDS wherever a replica fits")
    ap.add_argument("--records", type=int, default=1 << 20, help="learned_fcc32: training records")
    ap.add_argument("--epochs", type=int, default=4, help="learned_fcc32: training epochs")
    ap.add_argument("--tilt", type=float, default=0.125, help="learned_fcc32: 'const' cone tilt")
    ap.add_argument("--ess-warmup", type=int, default=500, help="learned_fcc32: sweeps before the trace")
    ap.add_argument("--start", default="ordered", choices=("ordered", "random"),
                    help="learned_fcc32: initial spins")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 "
              "(sharding replica ladders over GPUs is not built)",
              file=sys.stderr)
        sys.exit(2)
    st, wu = STEP_DEFAULTS.get(args.config, (10000, 1000))
    if args.steps is None:
        args.steps = st
    if args.warmup is None:
        args.warmup = wu

    if args.config == "batch_c3":
        return run_batch(args)
    if args.config == "learned_fcc32":
        return run_learned(args)
    if args.config == "copy_probe":
        return run_copy_probe(args)
    run_single(args, args.config, args.steps, args.warmup, with_cpu=not args.no_cpu)


if __name__ == "__main__":
    main()
