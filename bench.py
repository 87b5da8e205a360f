"""Headline benchmark: spin updates/s and effective samples/s of E of the device
Metropolis sweep on bcc 32^3 (65,536 spins; body-2 l<=2 to the 3rd shell, l=1
triangles and tetrahedra, isotropic SALC tensors).

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


def cpu_baseline(na, terms, dims, cfg, kT, budget_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    T = O.Tiled(na, terms, dims)
    nthreads = max(1, min(16, os.cpu_count() or 1))
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
    return {"value": T.n_active * n / dt, "unit": "spin updates/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{n} Metropolis sweeps of the reference CPU path (site_coeffs! + "
                      f"delta_energy, colour classes over {nthreads} threads) on the same "
                      f"{T.n_sites}-spin model, {dt:.1f} s"}


def _mean_err(x, nbins=20):
    x = np.asarray(x)
    b = x[: len(x) // nbins * nbins].reshape(nbins, -1).mean(1)
    return float(x.mean()), float(b.std(ddof=1) / math.sqrt(nbins))


def da_leg(args, S, H, gH, gst, beta, obs_plain, wall_plain):
    """Learned delayed acceptance against the plain sweep at the same kT: harvest
    (phi(e), phi(e'), exact dE) tuples from plain sweeps of the equilibrated chain,
    train the surrogate on the GPU, then time `steps` DA sweeps from the same state
    with E recorded every sweep."""
    from sce_amd import surrogate as SG

    sur = SG.Surrogate(gH, seed=424242)
    n_harvest = max(1, int(math.ceil(args.da_rows / H.n_active)))
    t0 = time.perf_counter()
    sur.harvest(gst, beta, n_harvest)
    t_harvest = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = sur.train_all(epochs=args.da_epochs, batch=args.da_batch, lr=args.da_lr)
    t_train = time.perf_counter() - t0
    cfg0, z0, e0 = gst.download()
    dst = S.GPUChainState(gH, S.ChainState(H, cfg0, gst.step), seed=777)
    SG.gpu_da_run_sweeps(dst, sur, None, beta, max(args.warmup // 4, 1), renorm_interval=1000)
    dst.reset_counters()
    t0 = time.perf_counter()
    obs = SG.gpu_da_run_sweeps(dst, sur, None, beta, args.steps, renorm_interval=1000, measure_interval=1)
    wall = time.perf_counter() - t0
    inf = dst.info()
    ess = S.effective_samples(obs[:, 0])
    ess_p = S.effective_samples(obs_plain[:, 0])
    e_da, de_da = _mean_err(obs[:, 0] / H.n_sites)
    e_pl, de_pl = _mean_err(obs_plain[:, 0] / H.n_sites)
    ess_s = ess / wall if math.isfinite(ess) else None
    ess_s_p = ess_p / wall_plain if math.isfinite(ess_p) else None
    return {
        "spin_updates_per_s": H.n_active * args.steps / wall,
        "ms_per_step": 1e3 * wall / args.steps,
        "effective_samples_per_s_E": ess_s,
        "ess_per_s_ratio_vs_plain": (ess_s / ess_s_p) if ess_s and ess_s_p else None,
        "tau_int_E": (args.steps / ess - 1) / 2 if math.isfinite(ess) and ess > 0 else None,
        "acceptance": inf.acc_metro / max(inf.att_metro, 1),
        "stage1_pass": inf.acc_stage1 / max(inf.att_metro, 1),
        "stage2_acceptance": inf.acc_metro / max(inf.acc_stage1, 1),
        "E_per_spin": [e_da, de_da], "E_per_spin_plain": [e_pl, de_pl],
        "E_diff_sigma": abs(e_da - e_pl) / math.hypot(de_da, de_pl) if de_da > 0 else None,
        "harvest_tuples": n_harvest * H.n_active, "harvest_s": t_harvest,
        "train_s": t_train, "train_loss_last_epoch": {str(a): float(l[-1]) for a, l in losses.items()},
        "surrogate": "per-sublattice MLP F->64->64->1 (tanh) + linear path, FP32; "
                     f"F = {[sur.info(a)['F'] for a in sur.active_sublattices()]}",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=3000)
    ap.add_argument("--workload", default="bcc32")
    ap.add_argument("--lanes", type=int, default=64)
    ap.add_argument("--kT", type=float, default=2.0e-3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--persistent", action="store_true",
                    help="one cooperative launch per renormalisation segment")
    ap.add_argument("--no-da", action="store_true", help="skip the delayed-acceptance leg")
    ap.add_argument("--da-rows", type=float, default=1e7, help="harvested training tuples")
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

    na, terms, dims = M.config_model(args.workload, seed=0)
    H = S.TiledHamiltonian(na, terms, dims)
    gH = S.GPUTiledHamiltonian(H)
    cfg = S.random_config(H.n_sites, np.random.default_rng(0))
    st = S.ChainState(H, cfg, 0.6)
    gst = S.GPUChainState(gH, st, seed=20261203)
    beta = 1.0 / args.kT
    gst.set_persistent(args.persistent)
    S.gpu_run_sweeps(gst, gH, None, beta, args.warmup, renorm_interval=1000, lanes=args.lanes)
    # kernel timing pass: HIP events around every colour-walk launch on the
    # chain's stream (sweeps launched one by one, not replayed as graphs)
    gst.set_timing(True)
    S.gpu_run_sweeps(gst, gH, None, beta, max(args.steps // 10, 20), renorm_interval=1000,
                     lanes=args.lanes, measure_interval=1)
    tot_ms_t, walk_ms_avg, walk_n = gst.timing()
    gst.set_timing(False)
    gst.reset_counters()
    # the timed region: `steps` sweeps with E and m recorded on the device every sweep
    t0 = time.perf_counter()
    obs = S.gpu_run_sweeps(gst, gH, None, beta, args.steps, renorm_interval=1000,
                           lanes=args.lanes, measure_interval=1)
    wall = time.perf_counter() - t0
    inf = gst.info()
    ups = H.n_active * args.steps / wall
    ess = S.effective_samples(obs[:, 0])
    b_site = algorithmic_bytes_per_update(H)
    n_timed = max(args.steps // 10, 20)
    walk_s = walk_ms_avg * walk_n / 1e3
    achieved = b_site * H.n_active * n_timed / walk_s / 1e9 if walk_s > 0 else float("nan")
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
        "data": "synthetic (seeded random spins, seeded N(0, 1 meV) isotropic SALC couplings)",
        "config": {"workload": f"{args.workload}: bcc {dims[0]}^3 x 2 atoms = {H.n_sites} spins, "
                               "body-2 l<=2 to 3rd shell + l=1 triangles + l=1 tetrahedra",
                   "kT_eV": args.kT, "lanes_per_site": args.lanes, "colours": H.n_colors},
        "effective_samples_per_s_E": ess / wall if math.isfinite(ess) else None,
        "tau_int_E": (args.steps / ess - 1) / 2 if math.isfinite(ess) and ess > 0 else None,
        "acceptance": inf.acc_metro / max(inf.att_metro, 1),
        "launch_mode": "persistent" if args.persistent else "hipGraph per sweep",
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_metro (colour walk)", "bytes_per_update": b_site,
                     "walk_launch_ms_avg": walk_ms_avg, "walk_launches": walk_n},
    }
    if not args.no_da:
        line["delayed_acceptance"] = da_leg(args, S, H, gH, gst, beta, obs, wall)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(na, terms, dims, cfg, args.kT)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
