"""Headline benchmark: spin updates/s and effective samples/s of E of the device
Metropolis sweep on bcc 32^3 (65,536 spins; body-2 l<=2 to the 3rd shell, l=1
triangles and tetrahedra, seeded synthetic isotropic tensors).

One "step" is one Metropolis sweep of the chain (every active site attempted
once), with the energy recorded on the device every sweep.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "scemontecarlo.jl_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_update(H) -> float:
    """Compulsory HBM bytes of one site update on the reference table layout:
    site_ptr pair + colour list (12), site_prog + site_col per pair entry (8),
    + site_col2 per triplet entry (12), site_prog + site_col + site_inst + the
    member share of the 20-byte instance record per general entry (8 + 9), own
    rows (8 nlm) + spin (24), dE/acc staging (12), accepted writes at p = 0.5."""
    s = 0 if H.n_active == 0 else int(H.color_sites[0])
    lo, hi = int(H.site_ptr[s - 1]), int(H.site_ptr[s])
    col = H.site_col[lo - 1:hi - 1]
    a_pair = int(np.sum(col > 0))
    a_trip = int(np.sum(col < 0))
    a_gen = int(np.sum(col == 0))
    nlm = H.nlm
    return (12 + 8 * a_pair + 12 * a_trip + 8 * a_gen + 9 * a_gen + 8 * nlm + 24 + 12
            + 0.5 * (8 * nlm + 36))


def fp64_flops_per_update(H) -> float:
    """FP64 flops of the walk of one site update: 3 per pair program entry, 4 per
    triplet entry, 6 per general (body-4) entry, as counted on the reference
    layout (product, weight, fused multiply-add)."""
    s = 0 if H.n_active == 0 else int(H.color_sites[0])
    flops = 0
    for j in range(int(H.site_ptr[s - 1]), int(H.site_ptr[s])):
        pid = int(H.site_prog[j - 1])
        n = int(H.sprog_ptr[pid]) - int(H.sprog_ptr[pid - 1])
        col = int(H.site_col[j - 1])
        flops += n * (3 if col > 0 else 4 if col < 0 else 6)
    return float(flops)


def cpu_baseline(na, terms, dims, cfg, kT, budget_s=4.0):
    """The reference CPU path (oracle restatement of metropolis_sweep!: per-site
    xoshiro256++ streams, colour classes split over threads with a barrier) on the
    same model, timed at 1, 4, 16 and nproc threads (the CPUs this process may run
    on) for about `budget_s` each; `value` is the fastest of them."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    T = O.Tiled(na, terms, dims)
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    counts = sorted({1, min(4, ncpu), min(16, ncpu), ncpu})
    by = {}
    detail = []
    for nthreads in counts:
        ch = O.CpuChain(T, cfg, 12345, 0.6)
        ch.metropolis_sweep(1.0 / kT, nthreads)     # warm
        t0 = time.perf_counter()
        n = 0
        while True:
            ch.metropolis_sweep(1.0 / kT, nthreads)
            n += 1
            if time.perf_counter() - t0 > budget_s or n >= 200:
                break
        dt = time.perf_counter() - t0
        by[str(nthreads)] = T.n_active * n / dt
        detail.append(f"{n} sweeps/{dt:.1f} s at {nthreads} thread(s)")
    best = max(counts, key=lambda k: by[str(k)])
    return {"value": by[str(best)], "unit": "spin updates/s", "cores": best, "kind": "port",
            "by_threads": by, "nproc": ncpu,
            "sample": f"Metropolis sweeps of the reference CPU path (site_coeffs! + delta_energy, "
                      f"colour classes over the threads) on the same {T.n_sites}-spin model: "
                      + "; ".join(detail)}


def measured_traffic(workload):
    """HBM bytes per walk launch from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a known-size copy)."""
    path = os.path.join(ROOT, "profiles", f"k_metro_{workload}_pmc.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    if "```json" not in text:
        return None
    blob = text.split("```json", 1)[1].split("```", 1)[0]
    try:
        return json.loads(blob).get("traffic_bytes_per_launch")
    except ValueError:
        return None


def train_surrogate(args, S, H, gH, gst, beta):
    """Harvest (phi(e), phi(e'), exact dE) tuples from plain sweeps of the
    equilibrated chain and train the surrogate on the GPU."""
    from sce_amd import surrogate as SG

    sur = SG.Surrogate(gH, seed=424242)
    n_harvest = max(1, int(math.ceil(args.da_rows / H.n_active)))
    t0 = time.perf_counter()
    sur.harvest(gst, beta, n_harvest)
    t_harvest = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = sur.train_all(epochs=args.da_epochs, batch=args.da_batch, lr=args.da_lr)
    t_train = time.perf_counter() - t0
    sur.set_int8(args.da_stage1 == "int8")
    return sur, {"harvest_tuples": n_harvest * H.n_active, "harvest_s": t_harvest, "train_s": t_train,
                 "train_loss_last_epoch": {str(a): float(l[-1]) for a, l in losses.items()}}


def ess_leg(args, S, H, gH, snapshot, step, beta):
    """`ess_sweeps` plain sweeps (graph-replayed, lanes as the timed region) from
    the equilibrated snapshot, E and m recorded every sweep: effective samples of
    E per second with tau_int(E) from the log-binning plateau, and Sokal's
    windowed estimate beside it."""
    pst = S.GPUChainState(gH, S.ChainState(H, snapshot, step), seed=778)
    pst.set_graph(not args.no_graph)
    S.gpu_run_sweeps(pst, gH, None, beta, 8, renorm_interval=1000, lanes=args.lanes, measure_interval=1)
    t0 = time.perf_counter()
    obs = S.gpu_run_sweeps(pst, gH, None, beta, args.ess_sweeps, renorm_interval=1000,
                           lanes=args.lanes, measure_interval=1)
    wall = time.perf_counter() - t0
    n = args.ess_sweeps
    ess = S.effective_samples(obs[:, 0])
    return {"sweeps": n, "wall_s": wall, "ms_per_step": 1e3 * wall / n,
            "spin_updates_per_s": H.n_active * n / wall,
            "effective_samples_per_s_E": ess / wall if math.isfinite(ess) else None,
            "tau_int_E": (n / ess - 1) / 2 if math.isfinite(ess) and ess > 0 else None,
            "tau_int_E_windowed": S.tau_int_windowed(obs[:, 0]),
            "obs": obs}


def da_leg(args, S, H, gH, sur, train_info, snapshot, step, beta, plain):
    """Delayed acceptance against the plain sweep at the same kT: `ess_sweeps`
    DA sweeps and as many plain sweeps (graph-replayed, lanes as the timed
    region), both from the equilibrated snapshot the plain timed region started
    from, with E and m recorded every sweep.  The ESS ratio rests on tau_int(E)
    estimates, whose spread over a few thousand sweeps is tens of percent, so
    this comparison runs longer than the timed region."""
    from sce_amd import surrogate as SG

    obs_plain, wall_plain = plain["obs"], plain["wall_s"]
    # the other stage-1 precision, timed over a shorter run from the same state
    other = "fp32" if args.da_stage1 == "int8" else "int8"
    sur.set_int8(other == "int8")
    ost = S.GPUChainState(gH, S.ChainState(H, snapshot, step), seed=777)
    ost.set_graph(not args.no_graph)
    SG.gpu_da_run_sweeps(ost, sur, None, beta, 8, renorm_interval=1000, measure_interval=1)
    ost.reset_counters()
    n_other = min(args.ess_sweeps, 4000)
    t0 = time.perf_counter()
    SG.gpu_da_run_sweeps(ost, sur, None, beta, n_other, renorm_interval=1000, measure_interval=1)
    w_other = time.perf_counter() - t0
    oi = ost.info()
    other_leg = {"sweeps": n_other, "ms_per_step": 1e3 * w_other / n_other,
                 "stage1_pass": oi.acc_stage1 / max(oi.att_metro, 1),
                 "stage2_acceptance": oi.acc_metro / max(oi.acc_stage1, 1)}
    sur.set_int8(args.da_stage1 == "int8")
    dst = S.GPUChainState(gH, S.ChainState(H, snapshot, step), seed=777)
    dst.set_graph(not args.no_graph)
    SG.gpu_da_run_sweeps(dst, sur, None, beta, 8, renorm_interval=1000, measure_interval=1)
    dst.reset_counters()
    t0 = time.perf_counter()
    obs = SG.gpu_da_run_sweeps(dst, sur, None, beta, args.ess_sweeps, renorm_interval=1000, measure_interval=1)
    wall = time.perf_counter() - t0
    inf = dst.info()
    ess = S.effective_samples(obs[:, 0])   # NaN when the timed series is too short for a plateau
    ess_p = S.effective_samples(obs_plain[:, 0])
    kT = 1.0 / beta
    sd = S.standard_stats(obs, kT, H.n_active)
    sp = S.standard_stats(obs_plain, kT, H.n_active)
    agree = {}
    for k in ("energy", "specific_heat", "absm", "susceptibility", "binder"):
        sig = math.hypot(sd[k].err, sp[k].err)
        agree[k] = {"da": sd[k].mean, "plain": sp[k].mean, "sigma": sig,
                    "diff_in_sigma": abs(sd[k].mean - sp[k].mean) / sig if sig > 0 else None}
    ess_s = ess / wall if math.isfinite(ess) else None
    ess_s_p = ess_p / wall_plain if math.isfinite(ess_p) else None
    n = args.ess_sweeps
    tw, tw_p = S.tau_int_windowed(obs[:, 0]), S.tau_int_windowed(obs_plain[:, 0])
    return {
        "sweeps": n,
        "spin_updates_per_s": H.n_active * n / wall,
        "ms_per_step": 1e3 * wall / n,
        "effective_samples_per_s_E": ess_s,
        "ess_per_s_ratio_vs_plain": (ess_s / ess_s_p) if ess_s and ess_s_p else None,
        # Peskun: with the same proposals DA accepts no more often than the plain
        # chain, so its tau_int is no smaller and its ESS/s ratio is at most the
        # ratio of sweep rates; a measured ratio above this is tau-estimate noise
        "ess_per_s_ratio_bound": wall_plain / wall,
        "tau_int_E": (n / ess - 1) / 2 if math.isfinite(ess) and ess > 0 else None,
        "plain_same_length": {"ms_per_step": 1e3 * wall_plain / n, "effective_samples_per_s_E": ess_s_p,
                              "tau_int_E": (n / ess_p - 1) / 2 if math.isfinite(ess_p) and ess_p > 0 else None,
                              "tau_int_E_windowed": tw_p},
        # cross-check with the autocorrelation summed to Sokal's window, whose
        # estimate of a short tau scatters far less than the plateau's
        "tau_int_E_windowed": tw,
        "ess_per_s_ratio_windowed": (wall_plain * (1 + 2 * tw_p)) / (wall * (1 + 2 * tw)),
        "acceptance": inf.acc_metro / max(inf.att_metro, 1),
        "stage1_pass": inf.acc_stage1 / max(inf.att_metro, 1),
        "stage2_acceptance": inf.acc_metro / max(inf.acc_stage1, 1),
        "observables_vs_plain": agree,
        **train_info,
        "surrogate": "per-sublattice MLP F->64->64->1 (tanh) + linear path, trained in FP32; "
                     f"F = {[sur.info(a)['F'] for a in sur.active_sublattices()]}",
        "stage1": args.da_stage1 + (" (int8 descriptor and weights, integer dot products on the matrix cores)"
                                    if args.da_stage1 == "int8" else " (FP32 fused multiply-add chains)"),
        f"stage1_{other}": other_leg,
    }


def da_kt_scan(args, S, H, gH, sur, snapshot, step, kts=(0.02, 0.025, 0.03, 0.035, 0.04), equil=2000,
               sweeps=8000):
    """Delayed acceptance against the plain sweep across kT at the frozen step
    adapted at the bench's kT (the surrogate as trained there): at each kT, `equil`
    plain sweeps from the equilibrated snapshot, then `sweeps` DA and `sweeps`
    plain sweeps from that same state (graph-replayed, E and m every sweep).
    Reports the stage-1 pass rate (the share of proposals walked exactly),
    tau_int(E) (log-binning plateau and Sokal's window), sweep times, ESS/s and
    the largest observable disagreement in sigma."""
    from sce_amd import surrogate as SG
    out = []
    for k, kT in enumerate(kts):
        beta = 1.0 / kT
        g0 = S.GPUChainState(gH, S.ChainState(H, snapshot, step), seed=3100 + k)
        S.gpu_run_sweeps(g0, gH, None, beta, equil, renorm_interval=1000, lanes=args.lanes)
        start = g0.download()[0]
        row = {"kT_eV": kT}
        obs = {}
        for name in ("plain", "da"):
            g = S.GPUChainState(gH, S.ChainState(H, start, step), seed=3200 + k)
            if name == "plain":
                S.gpu_run_sweeps(g, gH, None, beta, 8, renorm_interval=1000, lanes=args.lanes, measure_interval=1)
                g.reset_counters()
                t0 = time.perf_counter()
                o = S.gpu_run_sweeps(g, gH, None, beta, sweeps, renorm_interval=1000, lanes=args.lanes,
                                     measure_interval=1)
            else:
                SG.gpu_da_run_sweeps(g, sur, None, beta, 8, renorm_interval=1000, measure_interval=1)
                g.reset_counters()
                t0 = time.perf_counter()
                o = SG.gpu_da_run_sweeps(g, sur, None, beta, sweeps, renorm_interval=1000, measure_interval=1)
            wall = time.perf_counter() - t0
            inf = g.info()
            ess = S.effective_samples(o[:, 0])
            tw = S.tau_int_windowed(o[:, 0])
            obs[name] = o
            row[name] = {"ms_per_sweep": 1e3 * wall / sweeps, "acceptance": inf.acc_metro / max(inf.att_metro, 1),
                         "tau_int_E": (sweeps / ess - 1) / 2 if math.isfinite(ess) and ess > 0 else None,
                         "tau_int_E_windowed": tw,
                         "ess_per_s_E_windowed": sweeps / (1 + 2 * tw) / wall if math.isfinite(tw) else None}
            if name == "da":
                row[name]["stage1_pass"] = inf.acc_stage1 / max(inf.att_metro, 1)
                row[name]["stage2_acceptance"] = inf.acc_metro / max(inf.acc_stage1, 1)
        sd = S.standard_stats(obs["da"], kT, H.n_active)
        sp = S.standard_stats(obs["plain"], kT, H.n_active)
        row["max_observable_diff_sigma"] = max(
            abs(sd[k2].mean - sp[k2].mean) / math.hypot(sd[k2].err, sp[k2].err)
            for k2 in ("energy", "specific_heat", "absm", "susceptibility", "binder")
            if math.hypot(sd[k2].err, sp[k2].err) > 0)
        a, b = row["da"]["ess_per_s_E_windowed"], row["plain"]["ess_per_s_E_windowed"]
        row["ess_per_s_ratio_windowed"] = a / b if a and b else None
        row["exact_walk_bound"] = 1.0 / max(row["da"]["stage1_pass"], 1e-9)
        out.append(row)
    return {"kts_eV": list(kts), "step": step, "equilibration_sweeps": equil, "sweeps": sweeps, "by_kT": out,
            "note": "exact_walk_bound = 1 / stage-1 pass rate: the most a free surrogate could gain"}


def _batch_candidates(form):
    """Autotuner candidates of the batch legs' --batch-form: both launch forms,
    per-colour launches only, or the persistent launch only."""
    from sce_amd.batch import _BATCH_CANDIDATES
    return tuple(c for c in _BATCH_CANDIDATES if form == "auto" or c[1] == (form == "persistent"))


def pt_leg(S, M, sweeps_therm=200, sweeps_measure=500, graph=True, form="auto", lanes=None, share=None):
    """Batched parallel tempering: 64 replicas of fcc 12^3 (6912 spins, body-2 l<=2 to
    the 4th shell) on a geometric kT ladder spanning 0.2-2.0 J1, replica exchange
    on the device after every sweep.  The autotuner picks lanes per site and the
    launch form: one launch per colour for all replicas (replayed hipGraphs) or
    one persistent launch per segment with the exchange decided in-kernel.
    J1 = 3 x rms raw coefficient of the first-shell (1,1) bonds (the bond energy
    of a (1,1) template is 4 pi N1^2 c = 3c times a unit-norm tensor form)."""
    na, terms, dims = M.config_model("fcc12", seed=0)
    first = [t for t in terms if len(t.atoms) == 2 and list(t.ls) == [1, 1]
             and np.linalg.norm(np.asarray(M.BASIS["fcc"][t.atoms[1] - 1]) + np.asarray(t.shifts[1])
                                - np.asarray(M.BASIS["fcc"][t.atoms[0] - 1]) - np.asarray(t.shifts[0]))
             < M.shell_distances("fcc", 1)[0] + 1e-9]
    J1 = 3.0 * math.sqrt(np.mean([t.coef ** 2 for t in first]))
    R = 64
    kts = J1 * np.geomspace(0.2, 2.0, R)
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    from sce_amd.batch import GPUBatch, derive_seeds, _adapt, autotune_batch, set_walk_form
    init, exch, dev = derive_seeds(2026, R)

    def make():
        states = [S.ChainState(H, S.random_config(H.n_sites, np.random.default_rng(init[r])), 0.6)
                  for r in range(R)]
        b = GPUBatch([gH] * R, states, dev, kts, exch)
        b.set_graph(graph)
        return b

    tune = {}
    if lanes is None:
        # lanes per site and launch form (per-colour graphs or one persistent
        # launch per segment): scratch batches, before the run
        lanes, pers, walk = autotune_batch(make, _batch_candidates(form), sweeps=20, timings=tune)
    else:   # pinned by the caller: per-colour launches
        pers, walk = False, None
    b = make()
    b.set_persistent(pers)
    if walk is None:
        b.set_shared_walk(0 if share is None else share)
    else:
        set_walk_form(b, walk)
    walk_name = ({"": "one site", "s": "two replicas", "q": "site pairs"}[walk] if walk is not None else
                 {None: "library default", 0: "library default", 1: "one site"}.get(share, f"{share} replicas"))
    steps = [0.6] * R
    done = 0
    while done < sweeps_therm:            # untimed: exchange + step adaptation (gpu_pt.jl:115-140)
        n = min(50, sweeps_therm - done)
        b.run(n, lanes, 1000, 0, 1, done, sweeps_therm)
        done += n
        for k in range(R):
            inf = b.info(k)
            steps[k] = _adapt(steps[k], inf.acc_metro, inf.att_metro, 0.5)
            b.set_step(k, steps[k])
        b.reset_counters()
    b.renormalize()
    t0 = time.perf_counter()
    obs = b.run(sweeps_measure, lanes, 1000, 1, 1, 0, sweeps_measure)
    wall = time.perf_counter() - t0
    att, acc = b.swap_stats()
    sw = np.where(att > 0, acc / np.maximum(att, 1), np.nan)
    return {"workload": f"64 replicas x fcc 12^3 = {R * H.n_sites} spins, ladder {kts[0]:.2e}-{kts[-1]:.2e} eV "
                        f"(J1 = {J1:.3e} eV), exchange every sweep on the device",
            "spin_updates_per_s": R * H.n_active * sweeps_measure / wall,
            "bytes_per_update": algorithmic_bytes_per_update(H),
            "algorithmic_GBs": R * H.n_active * sweeps_measure / wall * algorithmic_bytes_per_update(H) / 1e9,
            "ms_per_sweep_all_replicas": 1e3 * wall / sweeps_measure, "sweeps": sweeps_measure,
            "lanes": lanes, "persistent": pers, "walk": walk_name, "measurements": int(obs.shape[1]),
            "autotune_ms_per_sweep": {k: 1e3 * v for k, v in tune.items()},
            "swap_acceptance_mean": float(np.nanmean(sw)), "swap_acceptance_min": float(np.nanmin(sw)),
            "note": "measurement phase: sweeps + device exchange + (E, m) of every replica every sweep; "
                    "autotune keys: lanes per site, 'p' = one persistent launch per segment, 's' = two "
                    "replicas per wavefront on one shared walk, 'q' = two same-sublattice sites of a replica "
                    "per wavefront (both: results those of 64 lanes per site)"}


def uneven_leg(S, M, sweeps=100, n_chains=256, graph=True, form="auto"):
    """Uneven batch: 256 chains, sc or bcc with N^3, N uniform in 6..20, each with
    its own seeded body-2 l<=2 (+ body-3 l=1 triangles on bcc) model, kT and a
    Zeeman term (2 mu_B moments, field 0-10 T along a random axis).  Chains run in
    three size classes (N <= 10, <= 15, <= 20), one batch per class, with lanes
    per site per class picked by the autotuner.  The timed run sweeps the
    three batches at once, each on its own stream (run_concurrent, per-colour
    launches: a persistent grid cannot share the GPU); each class is also
    timed alone.  With form "persistent" the classes run one after another."""
    from sce_amd.batch import GPUBatch, autotune_batch, run_concurrent, set_walk_form

    rng = np.random.default_rng(20261203)
    classes = {0: [], 1: [], 2: []}
    total_sites = 0
    for k in range(n_chains):
        lat = "sc" if rng.random() < 0.5 else "bcc"
        N = int(rng.integers(6, 21))
        na, terms = M.lattice_model(lat, pair_shells=2, pair_lmax=2, triangles=(lat == "bcc"), seed=1000 + k)
        B = rng.uniform(0.0, 10.0) * (lambda v: v / np.linalg.norm(v))(rng.standard_normal(3))
        H = S.TiledHamiltonian(na, terms, (N, N, N), magmoms=[2.0] * na, field=tuple(B))
        st = S.ChainState(H, S.random_config(H.n_sites, rng), 0.6)
        classes[0 if N <= 10 else 1 if N <= 15 else 2].append((H, st, int(rng.integers(1, 2**62)),
                                                                float(rng.uniform(0.005, 0.05))))
        total_sites += H.n_active
    out = {}
    t_seq = 0.0
    runs = []
    for cls, chains in classes.items():
        gHs = [S.GPUTiledHamiltonian(H) for H, _, _, _ in chains]

        def make():
            b = GPUBatch(gHs, [st for _, st, _, _ in chains], [sd for _, _, sd, _ in chains],
                         [kt for _, _, _, kt in chains])
            b.set_graph(graph)
            return b
        tune = {}
        cands = _batch_candidates("graph" if form == "auto" else form)
        lanes, pers, walk = autotune_batch(make, cands, sweeps=10, timings=tune)
        b = make()
        b.set_persistent(pers)
        set_walk_form(b, walk)
        b.run(5, lanes, 0)
        t0 = time.perf_counter()
        b.run(sweeps, lanes, 1000)
        dt = time.perf_counter() - t0
        t_seq += dt
        n_sites = sum(H.n_active for H, _, _, _ in chains)
        out[f"class_{cls}"] = {"chains": len(chains), "spins": n_sites, "lanes": lanes, "persistent": pers,
                               "walk": {"": "one site", "q": "site pairs"}.get(walk, walk),
                               "ms_per_sweep_alone": 1e3 * dt / sweeps,
                               "spin_updates_per_s_alone": n_sites * sweeps / dt,
                               "autotune_ms_per_sweep": {k: 1e3 * v for k, v in tune.items()}}
        runs.append((b, (sweeps, lanes, 1000), {"phase_done": sweeps}))
    # the timed run: every class at once, each batch on its own stream
    together = not any(c["persistent"] for c in out.values())
    t0 = time.perf_counter()
    if together:
        run_concurrent(runs)
    else:
        for b, a, k in runs:
            b.run(*a, **k)
    t_all = time.perf_counter() - t0
    return {"workload": f"{n_chains} chains, sc/bcc N^3 with N in 6..20, {total_sites} spins, per-chain model, "
                        "kT and Zeeman field; three size classes "
                        + ("sweeping at once on their own streams" if together else "one after another"),
            "spin_updates_per_s": total_sites * sweeps / t_all, "ms_per_sweep": 1e3 * t_all / sweeps,
            "spin_updates_per_s_classes_one_after_another": total_sites * sweeps / t_seq,
            "sweeps": sweeps, "classes": out}


def compound_leg(S, M, sweeps=10000, equil=1000, kts=(0.01, 0.02, 0.04, 0.08), graph=True):
    """Plain form on the compound sweep of run_mc (src/run.jl:143-150): bcc 16^3
    (8192 spins, body-2 l<=2 to the 3rd shell + l=1 triangles), each sweep one
    Metropolis sweep then one overrelaxation pass, `sweeps` measured sweeps at
    each of four fixed kT after `equil` sweeps with run_mc's step adaptation.
    Every compound sweep is one replayed hipGraph (measurement fused into the
    overrelaxation reduction, the next sweep's proposals made there too);
    Metropolis and overrelaxation attempts per second are reported apart.
    Each kT's chain is timed alone; then the four chains sweep `sweeps` more
    at once, each on its own stream (run_concurrent)."""
    from sce_amd.batch import _adapt, run_concurrent
    na, terms, dims = M.config_model("bcc16", seed=0)
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    out = []
    chains = []
    for k, kT in enumerate(kts):
        st = S.ChainState(H, S.random_config(H.n_sites, np.random.default_rng(100 + k)), 0.6)
        g = S.GPUChainState(gH, st, seed=9000 + k)
        g.set_graph(graph)
        beta = 1.0 / kT
        done = 0
        while done < equil:
            n = min(50, equil - done)
            g.reset_counters()
            S.gpu_run_sweeps(g, gH, None, beta, n, renorm_interval=0, or_per_metropolis=1)
            done += n
            if n == 50:
                inf = g.info()
                g.step = _adapt(g.step, inf.acc_metro, inf.att_metro, 0.5)
        S.gpu_renormalize(g)
        S.gpu_run_sweeps(g, gH, None, beta, 8, renorm_interval=1000, or_per_metropolis=1, measure_interval=1)
        g.reset_counters()
        t0 = time.perf_counter()
        obs = S.gpu_run_sweeps(g, gH, None, beta, sweeps, renorm_interval=1000, or_per_metropolis=1,
                               measure_interval=1)
        wall = time.perf_counter() - t0
        inf = g.info()
        stats = S.standard_stats(obs, kT, H.n_active)
        ess = S.effective_samples(obs[:, 0])
        row = {"kT_eV": kT, "ms_per_compound_sweep": 1e3 * wall / sweeps,
               "metropolis_attempts_per_s": inf.att_metro / wall, "or_attempts_per_s": inf.att_or / wall,
               "metropolis_acceptance": inf.acc_metro / max(inf.att_metro, 1),
               "or_acceptance": inf.acc_or / max(inf.att_or, 1), "step": g.step,
               "effective_samples_per_s_E": ess / wall if math.isfinite(ess) else None,
               "energy_per_site": stats["energy"].mean / H.n_active, "absm": stats["absm"].mean}
        if k == 0:   # the same sweeps launched directly (the form before graph capture)
            g.set_graph(False)
            t0 = time.perf_counter()
            S.gpu_run_sweeps(g, gH, None, beta, 1000, renorm_interval=1000, or_per_metropolis=1, measure_interval=1)
            row["ms_per_compound_sweep_direct_launches"] = 1e3 * (time.perf_counter() - t0) / 1000
            g.set_graph(graph)
        out.append(row)
        chains.append((g, beta))
    # the four kT at once: each chain on its own stream
    for g, _ in chains:
        g.reset_counters()
    runs = [(S.gpu_run_sweeps, (g, gH, None, beta, sweeps),
             {"renorm_interval": 1000, "or_per_metropolis": 1, "measure_interval": 1}) for g, beta in chains]
    t0 = time.perf_counter()
    run_concurrent(runs)
    wall = time.perf_counter() - t0
    infs = [g.info() for g, _ in chains]
    together = {"sweeps": sweeps, "ms_per_compound_sweep_of_all_chains": 1e3 * wall / sweeps,
                "metropolis_attempts_per_s": sum(i.att_metro for i in infs) / wall,
                "or_attempts_per_s": sum(i.att_or for i in infs) / wall}
    return {"workload": f"bcc16: bcc 16^3 x 2 atoms = {H.n_sites} spins, body-2 l<=2 to 3rd shell + l=1 "
                        "triangles; compound sweep = 1 Metropolis + 1 overrelaxation pass",
            "sweeps_per_kT": sweeps, "equilibration_sweeps": equil, "colours": H.n_colors, "by_kT": out,
            "all_kT_at_once": together}


def setup_leg(S, sweeps=200, graph=True):
    """The Nd2Fe14B cell read from its setup file (tests/golden/nd2fe14b.toml, the
    data of bench/assets/nd2fe14b.toml), modelled as the reference's triplet-heavy
    fixture asks (bench/fixtures.jl:90-100: nbody 3, 3.5 A cutoff, l <= 2 on Nd and
    Fe, boron inactive) by the seeded generator, at 4^3 and 8^3 cells, kT 0.025 eV
    (fixtures.jl BENCH_KT), lanes picked per size.  The generator's isotropic basis
    has fewer terms per cell than a fitted SCEFitting basis on the same cell, so
    the terms and the device walk's entries per site are reported beside the
    sweep time."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "nd2fe14b.toml")
    st = S.read_setup(path)
    na, terms = S.setup_model(st, nbody=3, cutoff=3.5, lmax=[2] * 8 + [0], seed=13)
    out = {"workload": "Nd2Fe14B 68-atom cell, nbody 3, cutoff 3.5 A, isotropic l <= 2, seeded 0.01 eV "
                       "coefficients", "terms_per_cell": len(terms),
           "body3_terms_per_cell": sum(1 for t in terms if len(t.atoms) == 3)}
    rng = np.random.default_rng(13)
    for n in (4, 8):
        H = S.TiledHamiltonian(na, terms, (n, n, n))
        gH = S.GPUTiledHamiltonian(H)
        cfg = S.random_config(H.n_sites, rng)
        best = None
        for lanes in (16, 32, 64, 128, 256):
            g = S.GPUChainState(gH, S.ChainState(H, cfg, 0.6), seed=5)
            g.set_graph(graph)
            S.gpu_run_sweeps(g, gH, None, 1 / 0.025, 20, renorm_interval=0, lanes=lanes)
            t0 = time.perf_counter()
            S.gpu_run_sweeps(g, gH, None, 1 / 0.025, sweeps, renorm_interval=1000, lanes=lanes)
            ms = 1e3 * (time.perf_counter() - t0) / sweeps
            if best is None or ms < best[0]:
                best = (ms, lanes)
            del g
        # the device walk holds one program per active sublattice (one site per cell)
        entries = H.sent_w.size * n ** 3
        out[f"{n}^3"] = {"sites": H.n_sites, "active": H.n_active, "colors": H.n_colors,
                         "walk_entries_per_active_site": entries / max(H.n_active, 1),
                         "ms_per_sweep": best[0], "lanes": best[1],
                         "spin_updates_per_s": H.n_active / (best[0] * 1e-3),
                         "walk_entries_per_s": entries / (best[0] * 1e-3), "sweeps": sweeps}
        del gH
    return out


def device_smoke(S, H, gH):
    """On-device gates before timing (bench/bench_gpu.jl:112-132): two chains with one
    seed must agree bitwise after 20 sweeps, and the incremental energy must stay
    within 1e-8 max(1, |E|) of a full recompute.  Aborts the run otherwise."""
    cfg = S.random_config(H.n_sites, np.random.default_rng(99))
    a = S.GPUChainState(gH, S.ChainState(H, cfg, 0.6), seed=5)
    b = S.GPUChainState(gH, S.ChainState(H, cfg, 0.6), seed=5)
    S.gpu_run_sweeps(a, gH, None, 1 / 0.04, 20, renorm_interval=0)
    S.gpu_run_sweeps(b, gH, None, 1 / 0.04, 20, renorm_interval=0)
    ca, za, Ea = a.download()
    cb, zb, Eb = b.download()
    E_full = a.total_energy()
    if not (np.array_equal(ca, cb) and np.array_equal(za, zb) and Ea == Eb):
        sys.exit("device smoke: repeated runs differ")
    if abs(Ea - E_full) > 1e-8 * max(1.0, abs(E_full)):
        sys.exit(f"device smoke: drift {abs(Ea - E_full):.3e} above 1e-8 max(1, |E|)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=500,
                    help="untimed sweeps at the frozen step, run right before the timed region")
    ap.add_argument("--equil", type=int, default=4000,
                    help="fixed equilibration with run_mc's step adaptation (every 50 sweeps toward "
                         "acceptance 0.5), run whatever --warmup is, before every measured leg")
    ap.add_argument("--workload", default="bcc32")
    ap.add_argument("--lanes", type=int, default=64)
    ap.add_argument("--kT", type=float, default=0.04,
                    help="model units (eV); at 0.04 the 32^3 chain equilibrates (tau_int(E) ~ 4 "
                         "sweeps), at <= 0.02 the random-sign couplings freeze into a glass")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--persistent", action="store_true",
                    help="one cooperative launch per renormalisation segment")
    ap.add_argument("--no-da", action="store_true", help="skip the delayed-acceptance leg")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched legs (64-replica PT, 256-chain uneven batch)")
    ap.add_argument("--no-compound", action="store_true",
                    help="skip the bcc 16^3 Metropolis + overrelaxation leg (4 kT x 10k sweeps)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched legs (64-replica PT, 256-chain uneven batch) and print their JSON")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the single-chain autotuner report (its persistent candidate is a cooperative launch)")
    ap.add_argument("--compound-only", action="store_true",
                    help="run only the bcc 16^3 Metropolis + overrelaxation leg and print its JSON")
    ap.add_argument("--pt-only", action="store_true",
                    help="run only the 64-replica fcc 12^3 ladder, at --pt-lanes (autotuned when not given) and "
                         "--pt-shared-walk, for --pt-sweeps thermalisation and measured sweeps, and print its JSON")
    ap.add_argument("--pt-lanes", type=int, default=None)
    ap.add_argument("--pt-shared-walk", type=int, default=None,
                    help="replicas per shared walk at 64 lanes (0: the library default, 1: off, 2, 4)")
    ap.add_argument("--pt-sweeps", type=int, default=None)
    ap.add_argument("--no-setup-model", action="store_true",
                    help="skip the Nd2Fe14B setup-file leg (nbody 3 at 4^3 and 8^3 cells)")
    ap.add_argument("--setup-model-only", action="store_true",
                    help="run only the Nd2Fe14B setup-file leg and print its JSON")
    ap.add_argument("--batch-form", choices=("auto", "graph", "persistent"), default="auto",
                    help="launch forms the batch legs' autotuner may pick: per-colour launches (graph), one "
                         "persistent launch per segment, or either (auto)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every sweep's kernels directly instead of replaying hipGraphs")
    ap.add_argument("--da-rows", type=float, default=1e7, help="harvested training tuples")
    ap.add_argument("--no-da-scan", action="store_true",
                    help="skip the delayed-acceptance kT scan (5 kT x (2k + 2 x 8k) sweeps)")
    ap.add_argument("--ess-sweeps", type=int, default=20000,
                    help="sweeps of each chain in the delayed-acceptance vs plain comparison")
    ap.add_argument("--da-stage1", choices=("int8", "fp32"), default="int8",
                    help="precision of the surrogate stage of the delayed-acceptance legs (the other one "
                         "is timed beside it)")
    ap.add_argument("--da-epochs", type=int, default=2)
    ap.add_argument("--da-batch", type=int, default=4096)
    ap.add_argument("--da-lr", type=float, default=1e-3)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES five times (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": 4 * n, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (every box has "
              "one MI355X; replicas shard across GPUs by design only)", file=sys.stderr)
        sys.exit(2)

    import sce_amd as S
    from sce_amd import models as M

    if args.batch_only:
        print(json.dumps({"batched": {"parallel_tempering": pt_leg(S, M, graph=not args.no_graph, form=args.batch_form),
                                      "uneven_batch": uneven_leg(S, M, graph=not args.no_graph,
                                                                 form=args.batch_form)}}),
              flush=True)
        return
    if args.pt_only:
        n = args.pt_sweeps
        kw = {} if n is None else {"sweeps_therm": n, "sweeps_measure": n}
        print(json.dumps({"parallel_tempering": pt_leg(S, M, graph=not args.no_graph, form=args.batch_form,
                                                       lanes=args.pt_lanes, share=args.pt_shared_walk, **kw)}),
              flush=True)
        return
    if args.compound_only:
        print(json.dumps({"compound": compound_leg(S, M, graph=not args.no_graph)}), flush=True)
        return
    if args.setup_model_only:
        print(json.dumps({"setup_model": setup_leg(S, graph=not args.no_graph)}), flush=True)
        return
    na, terms, dims = M.config_model(args.workload, seed=0)
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    device_smoke(S, H, gH)
    cfg = S.random_config(H.n_sites, np.random.default_rng(0))
    st = S.ChainState(H, cfg, 0.6)
    gst = S.GPUChainState(gH, st, seed=20261203)
    beta = 1.0 / args.kT
    gst.set_persistent(args.persistent)
    gst.set_graph(not args.no_graph)
    # fixed equilibration with run_mc's step adaptation (src/run.jl:172: every 50
    # sweeps toward acceptance 0.5), independent of --warmup; then --warmup sweeps
    # at the frozen step, which stays frozen for everything measured
    from sce_amd.batch import _adapt
    done = 0
    while done < args.equil:
        n = min(50, args.equil - done)
        gst.reset_counters()
        S.gpu_run_sweeps(gst, gH, None, beta, n, renorm_interval=0, lanes=args.lanes)
        done += n
        if n == 50:
            inf = gst.info()
            gst.step = _adapt(gst.step, inf.acc_metro, inf.att_metro, 0.5)
    S.gpu_renormalize(gst)
    sur = train_info = None
    if not args.no_da:   # harvesting is plain sweeps: it continues the equilibration
        sur, train_info = train_surrogate(args, S, H, gH, gst, beta)
    # untimed: a few measured sweeps, so that the graph variants the timed
    # region replays (measured, chained) are captured before it starts
    S.gpu_run_sweeps(gst, gH, None, beta, 8, renorm_interval=1000, lanes=args.lanes, measure_interval=1)
    snapshot = gst.download()[0]
    # the autotuner's view of this chain (reported; the timed region keeps the
    # pinned --lanes and launch mode, which are part of the determinism scope)
    tl = tp = tt = None
    if not args.no_autotune:
        tl, tp, tt = S.autotune_chain(gH, S.ChainState(H, snapshot, gst.step), beta, sweeps=200)
    # kernel timing: the colour walks of 20 passes launched back to back between
    # one pair of HIP events on the chain's stream (state restored afterwards)
    walk_ms_avg, walk_n = gst.time_walk(args.lanes, 20)
    # --warmup untimed sweeps at the frozen step, as the same kind of call as the
    # timed region and right before it
    if args.warmup > 0:
        S.gpu_run_sweeps(gst, gH, None, beta, args.warmup, renorm_interval=1000, lanes=args.lanes,
                         measure_interval=1)
    gst.reset_counters()
    # the timed region: `steps` sweeps with E and m recorded on the device every sweep
    t0 = time.perf_counter()
    obs = S.gpu_run_sweeps(gst, gH, None, beta, args.steps, renorm_interval=1000,
                           lanes=args.lanes, measure_interval=1)
    wall = time.perf_counter() - t0
    inf = gst.info()
    ups = H.n_active * args.steps / wall
    ess = S.effective_samples(obs[:, 0])   # NaN when the timed series is too short for a plateau
    b_site = algorithmic_bytes_per_update(H)
    b_launch = b_site * H.n_active / H.n_colors
    f_site = fp64_flops_per_update(H)
    fp64_tflops = f_site * H.n_active / H.n_colors / (walk_ms_avg * 1e-3) / 1e12 if walk_ms_avg > 0 else float("nan")
    achieved = b_launch / (walk_ms_avg * 1e-3) / 1e9 if walk_ms_avg > 0 else float("nan")
    line = {
        "metric": "spin updates/s + effective samples/s of E at 32^3 (plain device sweep)",
        "value": ups,
        "unit": "spin updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random spins; seeded N(0, 1 meV) coefficients on isotropic tensors: "
                "identity (l, l) pair blocks, Levi-Civita triangles, three-pairing tetrahedra)",
        "config": {"workload": f"{args.workload}: bcc {dims[0]}^3 x 2 atoms = {H.n_sites} spins, "
                               "body-2 l<=2 to 3rd shell + l=1 triangles + l=1 tetrahedra",
                   "kT_eV": args.kT, "lanes_per_site": args.lanes, "colours": H.n_colors,
                   "step": gst.step},
        "timed_region_effective_samples_per_s_E": ess / wall if math.isfinite(ess) else None,
        "timed_region_tau_int_E": (args.steps / ess - 1) / 2 if math.isfinite(ess) and ess > 0 else None,
        "acceptance": inf.acc_metro / max(inf.att_metro, 1),
        "autotune": {"seconds_per_sweep": tt, "fastest": {"lanes": tl, "persistent": tp},
                     "note": "200 measured sweeps of a throwaway copy of the chain per candidate "
                             "(lanes per site; p = dependency-driven persistent sweep)"},
        "launch_mode": ("persistent" if args.persistent else
                        "direct launches" if args.no_graph else "hipGraph per sweep"),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload),
                     "kernel": "k_metro (colour walk)", "bytes_per_update": b_site,
                     "bytes_per_launch": b_launch,
                     "walk_launch_ms_avg": walk_ms_avg, "walk_launches": walk_n,
                     "duration_source": "HIP events around 20 back-to-back passes of the colour "
                                        "walks on the chain's stream, just before the timed region",
                     "fp64": {"flops_per_update": f_site, "achieved": fp64_tflops, "peak": 78.6,
                              "unit": "TFLOP/s", "frac": fp64_tflops / 78.6},
                     # the whole sweep (16 walks, the reduction and the
                     # launch gaps) over the timed region, beside the kernel
                     "sweep_level": {"achieved": b_site * H.n_active / (wall / args.steps) / 1e9,
                                     "frac": b_site * H.n_active / (wall / args.steps) / 1e9 / HBM_PEAK_GBS,
                                     "unit": "GB/s"}},
    }
    # the all-site energy gradient (gpu_energy_gradient! twin) on the chain's rows
    grad_ms = S.gpu_time_gradient(gst, S.GRADIENT_LANES, 20)
    adj_bytes = 16 * int(H.site_prog.size)
    line["energy_gradient"] = {
        "kernel": f"k_grad_walk (all sites, {S.GRADIENT_LANES} strided shares per site)",
        "ms_per_call": grad_ms, "site_gradients_per_s": H.n_sites / (grad_ms * 1e-3),
        "adjacency_bytes_per_call": adj_bytes,
        "adjacency_GBs": adj_bytes / (grad_ms * 1e-3) / 1e9,
        "duration_source": "HIP events around 20 back-to-back launches on the chain's stream"}
    # effective samples of E per second: from a plain leg long enough for a
    # log-binning plateau (the timed region's own estimate is reported above,
    # or null when it is too short to have one)
    ess_leg_out = ess_leg(args, S, H, gH, snapshot, gst.step, beta)
    line["effective_samples_per_s_E"] = ess_leg_out["effective_samples_per_s_E"]
    line["tau_int_E"] = ess_leg_out["tau_int_E"]
    line["ess_leg"] = {k: v for k, v in ess_leg_out.items() if k != "obs"}
    if not args.no_da:
        line["delayed_acceptance"] = da_leg(args, S, H, gH, sur, train_info, snapshot, gst.step, beta,
                                            plain=ess_leg_out)
        if not args.no_da_scan:
            line["delayed_acceptance"]["kT_scan"] = da_kt_scan(args, S, H, gH, sur, snapshot, gst.step)
    if not args.no_batch:
        line["batched"] = {"parallel_tempering": pt_leg(S, M, graph=not args.no_graph, form=args.batch_form),
                           "uneven_batch": uneven_leg(S, M, graph=not args.no_graph, form=args.batch_form)}
    if not args.no_compound:
        line["compound"] = compound_leg(S, M, graph=not args.no_graph)
    if not args.no_setup_model:
        line["setup_model"] = setup_leg(S, graph=not args.no_graph)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(na, terms, dims, cfg, args.kT)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
