#!/usr/bin/env python3
"""Benchmark: Felsenstein-pruning CLV updates/s on one MI355X.

Workload (the headline configuration): GTR+G4 DNA, 128-taxon Yule tree with
Exp(mean 0.05) branch lengths, 100,000 sites, exchangeabilities ~
Dirichlet(1x6), base frequencies ~ Dirichlet(5x4), gamma shape 0.5 with 4
equal-weight categories, 5% of cells gap/'N'.  Partials stored in FP32 with
power-of-two scaling; arithmetic in FP64.

A step is one full likelihood evaluation (phylo_loglik: the pruning launch
over every internal node and site, the compensated total written to mapped
host memory, and the wait for it) with all inputs resident on the GPU.  The
default warm-up (300 evaluations) lets the kernel time settle: over the first
100 or so evaluations of a fresh process it falls by about 6 %.
One update = one (site, internal node, rate category) partial:
(n-1) * m * C = 127 * 100000 * 4 = 5.08e7 per step.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Bytes per update the pruning pass must move, FP32 partials, one int32
# scaler per (node, site) shared by the categories, 1-byte tip codes
# (per site: C*S*4*(n-2) child reads + n tip bytes + C*S*4*(n-1) parent
# writes + 4*(n-1) scaler writes + 4*(n-2) scaler reads, plus ~72 B root
# reduction, divided by the (n-1)*C updates per site).
def alg_bytes_per_update(n, S=4, C=4):
    per_site = C * S * 4 * (n - 2) + n + C * S * 4 * (n - 1) + 4 * (n - 1) + 4 * (n - 2) + 72
    return per_site / ((n - 1) * C)


HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
FP64_VALU_PEAK_TFLOPS = 78.6     # MI355X FP64 vector peak (equal to its dense FP64 matrix peak)


def flops_per_update(S):
    """FP64 flops of one (site, node, category) update: two P x L products
    (2 S^2 multiply-adds each) and the elementwise product of the children."""
    return 2 * 2 * S * S + S


def host_cores():
    """CPUs this process may run on (its affinity mask), capped at the
    CPU share the environment grants a job (OMP_NUM_THREADS when set): the
    count of concurrent CPU shards the baseline runs."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(share) if share.isdigit() and int(share) > 0 else avail
    return max(1, min(avail, cap)), avail


def cpu_baseline(wl, budget_s=20.0):
    """Reference CPU path timed on this host: the compiled phylommand simpleML
    (oracle/_ref/ref_harness, built with -O2 from the reference sources) on a
    bounded slice of the same workload; falls back to the oracle's restatement
    of the same cost model (P rebuilt per site and edge by squaring)."""
    from oracle import oracle as O  # checker only: the CPU baseline leg
    from phylommand_amd import synth
    t = wl.tree
    n = t.n_tips
    harness = O.REF_HARNESS
    m_sample = 1200
    cores, avail = host_cores()
    if os.path.exists(harness) and os.access(harness, os.X_OK):
        with tempfile.TemporaryDirectory() as d:
            open(os.path.join(d, "t.nwk"), "w").write(t.newick() + "\n")
            cmds = []
            for k in range(cores):
                # independent site shards: the reference is single-threaded and sites are independent
                masks = wl.masks[:, k * m_sample:(k + 1) * m_sample]
                fa = os.path.join(d, f"a{k}.fst")
                open(fa, "w").write(synth.fasta_text(t.names, masks, synth.DNA_LETTERS, 4))
                # the reference's sub_model cannot hold 6 distinct rates (heap overflow at
                # sub_model.cpp:87), so the timing runs a 1-rate reversible model: the cost
                # of a site-edge does not depend on the rate values.
                cmds.append([harness, "time", os.path.join(d, "t.nwk"), fa, "dna", "1", "0,0,0,0,0,0", "1.0",
                             ",".join("%.6f" % x for x in wl.pi[:3]), "3"])

            def secs(out):
                kv = dict(line.split(" ", 1) for line in out.strip().split("\n") if " " in line)
                return float(kv["seconds"])

            t0 = time.time()
            one = secs(subprocess.run(cmds[0], capture_output=True, text=True, timeout=600).stdout)
            procs = [subprocess.Popen(c, stdout=subprocess.PIPE, text=True) for c in cmds]
            many = [secs(p.communicate(timeout=600)[0]) for p in procs]
            wall = time.time() - t0
            updates = (n - 1) * m_sample
            rate_one = updates / one
            rate_all = sum(updates / s for s in many)
            return {"value": rate_all, "unit": "updates/s", "cores": cores, "kind": "reference",
                    "single_core": rate_one,
                    "cores_available": avail,
                    "sample": f"phylommand simpleML::calculate_log_likelihood compiled -O2 from the reference "
                              f"sources, {n} taxa x {m_sample} sites per core, {cores} concurrent site shards "
                              f"({avail} CPUs in the affinity mask, OMP_NUM_THREADS="
                              f"{os.environ.get('OMP_NUM_THREADS', 'unset')}; 1 core: {rate_one:.3g} updates/s); "
                              f"1 rate category, the reference has no gamma and its per-category cost is this; "
                              f"best of 3 calls per shard; {wall:.1f}s wall"}
    masks = wl.masks[:, :m_sample]
    # restatement of the reference's cost model
    mb = np.zeros((t.n_nodes, m_sample), np.uint32)
    mb[:n] = masks
    p = O.Problem(t.left, t.right, t.root, t.blen, mb, 4, wl.Q, wl.pi, pi=wl.pi, builder=O.REFCOMPAT)
    _, sec = p.time_reference_cost()
    return {"value": (n - 1) * m_sample / sec, "unit": "updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle restatement of simpleML's cost model, {n} taxa x {m_sample} sites, 1 category"}


_FAIR_CHILD = r"""
import sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import oracle as O
d = np.load(sys.argv[2])
k, ms = int(sys.argv[3]), int(sys.argv[4])
mb = np.ascontiguousarray(d["masks_by_node"][:, k * ms:(k + 1) * ms])
p = O.Problem(d["left"], d["right"], int(d["root"]), d["blen"], mb, 4, d["Q"], d["pi"], pi=d["pi"],
              cat_rates=d["cat_rates"], cat_w=d["cat_w"])
best = 1e30
for _ in range(3):
    t0 = time.perf_counter()
    p.loglik()
    best = min(best, time.perf_counter() - t0)
print(best)
"""


def fair_cpu(wl, cores, m_shard=20_000):
    """A CPU path without the reference's per-site P recomputation: the
    oracle's FP64 pruning (P built once per edge and category, -O2 C) with
    gamma categories, one process per core on independent site shards."""
    t = wl.tree
    n = t.n_tips
    m_shard = min(m_shard, wl.masks.shape[1] // cores)
    mb = np.zeros((t.n_nodes, m_shard * cores), np.uint32)
    mb[:n] = wl.masks[:, :m_shard * cores]
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "wl.npz")
        np.savez(f, masks_by_node=mb, left=t.left, right=t.right, root=t.root, blen=t.blen, Q=wl.Q, pi=wl.pi,
                 cat_rates=wl.cat_rates, cat_w=wl.cat_weights)
        t0 = time.time()
        procs = [subprocess.Popen([sys.executable, "-c", _FAIR_CHILD, ROOT, f, str(k), str(m_shard)],
                                  stdout=subprocess.PIPE, text=True) for k in range(cores)]
        secs = [float(p.communicate(timeout=600)[0].strip().split()[-1]) for p in procs]
        wall = time.time() - t0
    C = len(wl.cat_rates)
    per = (n - 1) * m_shard * C
    return {"value": sum(per / s for s in secs), "unit": "updates/s", "cores": cores,
            "single_process": per / max(secs),
            "sample": f"oracle FP64 pruning (P once per edge and category, gcc -O2), {n} taxa x {m_shard} sites "
                      f"x {C} categories per process, {cores} concurrent site shards, best of 3; {wall:.1f}s wall"}


def copy_bandwidth(mib=1024, reps=20):
    """Device-to-device copy rate (bytes read + written per second) of a
    plain 1 GiB copy on this GPU, timed with events on torch's stream."""
    import torch
    a = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / reps
    return 2 * a.numel() * 4 / sec / 1e9


def copy_bandwidth_subprocess():
    """copy_bandwidth in a child process that loads only torch's HIP runtime."""
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-probe", "1024", "--copy-time"],
                       capture_output=True, text=True, timeout=300)
    return float(json.loads(r.stdout.strip().splitlines()[-1])["copy_gbs"])


def profiled_traffic(name):
    """HBM bytes per launch of a kernel on its bench workload, from the
    FETCH_SIZE / WRITE_SIZE counter passes summarised in profiles/<name>.md
    (scaled by a copy kernel of known size); None if not profiled."""
    tf = os.path.join(ROOT, "profiles", name + "_traffic.json")
    try:
        return float(json.load(open(tf))["bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def measure_protein(P, synth, steps, warmup, m=20_000):
    """The 20-state walker on its workload: 64-taxon Yule tree, m sites, 190
    exchangeabilities ~ Gamma(1,1), frequencies ~ Dirichlet(2x20), gamma shape
    1.0 (4 categories), 3% gaps.  Returns the lnL, the wall time per step and
    the mean kernel time (HIP events on the library's stream) with the roofline
    on the FP64 matrix cores."""
    wl = synth.gtr_gamma_workload(n_tips=64, m=m, S=20, C=4, alpha=1.0, gap_frac=0.03, seed=4,
                                  rate_dist="gamma", freq_conc=2.0, mean_blen=0.05)
    t = wl.tree
    ctx = P.Context(0, P.PREC_FP32)
    ctx.set_tips(wl.masks, 20)
    ctx.set_topology(t.left, t.right, t.root)
    ctx.set_edge_lengths(t.blen)
    ctx.set_model(wl.Q, wl.pi, wl.pi, P.P_EIGEN)
    ctx.set_categories(wl.cat_rates, wl.cat_weights)
    for _ in range(warmup + 1):
        ctx.loglik()
    ctx.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        lnl = ctx.loglik()
        kms.append(ctx.last_kernel_ms())
    ctx.synchronize()
    el = time.perf_counter() - t0
    ctx.close()
    n, C, S = 64, 4, 20
    updates = (n - 1) * m * C
    ks = float(np.mean(kms)) * 1e-3
    # Matrix-core work: a stored or carried child's P x L is an FP64 MFMA chain
    # (2 S^2 flops per site and category); a tip child's is a gather of P
    # columns (no MFMA).  Counted over the sites, not the padding.
    internal_children = sum(int(c >= n) for v in range(t.n_nodes) for c in (t.left[v], t.right[v]) if c >= 0)
    mfma_flops = internal_children * m * C * 2 * S * S
    per_site = C * S * 4 * (n - 2) + n + C * S * 4 * (n - 1) + 4 * (n - 1) + 4 * (n - 2)
    bpu = per_site / ((n - 1) * C)
    traffic = profiled_traffic("prune_mfma_protein20_g4_64x20k") if m == 20_000 else None
    tf = mfma_flops / ks / 1e12
    return {"lnL": lnl, "updates": updates, "ms_per_step": el / steps * 1e3, "kernel_ms": ks * 1e3,
            "value": updates * steps / el,
            "roofline": {"bound": "mfma", "achieved": tf, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tf / FP64_MFMA_PEAK_TFLOPS, "traffic": traffic,
                         "kernel": "prune_mfma_kernel<float,4,5,2,true>", "kernel_ms": ks * 1e3,
                         "mfma_flops_per_step": mfma_flops,
                         "useful_flops_per_update": flops_per_update(S), "alg_bytes_per_update": bpu,
                         "hbm_gbs": traffic / ks / 1e9 if traffic else None,
                         "hbm_frac": traffic / ks / 1e9 / HBM_PEAK_GBS if traffic else None}}


def bench_protein(args, P, synth):
    """20-state protein, 64-taxon Yule tree, 20,000 sites, 190 exchangeabilities
    ~ Gamma(1,1), frequencies ~ Dirichlet(2x20), gamma shape 1.0 (4 categories),
    3% gaps."""
    m = args.sites or 20_000
    r = measure_protein(P, synth, args.steps, args.warmup, m)
    return {"metric": "CLV updates/s (site*node*cat)", "value": r["value"], "unit": "updates/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded Yule tree, 20-state simulation)",
            "config": {"workload": f"20-state protein + G4, 64 taxa Yule, {m} sites, 3% gaps", "n_taxa": 64,
                       "n_sites": m, "n_states": 20, "n_cat": 4, "clv_storage": "f32", "lnL": r["lnL"]},
            "roofline": r["roofline"]}


def bench_nni(args, P, synth):
    """Full NNI neighbourhood (2*(n-3) = 250 variants) of the 128-taxon x 100k GTR+G4 tree."""
    wl = synth.gtr_gamma_workload(n_tips=128, m=100_000, mean_blen=0.05, alpha=0.5, C=4, gap_frac=0.05, seed=2)
    t = wl.tree
    ctx = P.Context(0, P.PREC_FP32)
    ctx.set_tips(wl.masks, 4)
    ctx.set_topology(t.left, t.right, t.root)
    ctx.set_edge_lengths(t.blen)
    ctx.set_model(wl.Q, wl.pi, wl.pi, P.P_EIGEN)
    ctx.set_categories(wl.cat_rates, wl.cat_weights)
    ctx.loglik()
    for _ in range(args.warmup):
        ctx.nni_score(1)
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        base, var = ctx.nni_score(1)
        kms.append(ctx.last_kernel_ms())
    el = time.perf_counter() - t0
    ops = ctx.nni_ops()
    updates = ops * 100_000 * 4
    ks = float(np.mean(kms)) * 1e-3
    full_equiv = len(var) * 127 * 100_000 * 4
    return {"metric": "CLV updates/s (site*node*cat)", "value": updates * args.steps / el, "unit": "updates/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": "full NNI neighbourhood of GTR+G4 128 taxa x 100000 sites",
                                            "variants": len(var), "recompute_ops": ops,
                                            "full_reevaluation_updates": full_equiv,
                                            "variants_per_s": len(var) * args.steps / el},
            "kernel_ms": ks * 1e3}


def bench_batch(args, P, synth, n_inst=512):
    """512 independent GTR+G4 alignments, taxa ~ U{8..256}, sites log-uniform
    [500, 50000], per-instance Dirichlet parameters, each with its full NNI
    neighbourhood, scored by the persistent batch launches replayed from one
    captured HIP graph."""
    t0 = time.time()
    wls = synth.batch_workloads(n_inst, workers=host_cores()[0])
    print(f"[bench] batch: {n_inst} instances simulated ({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)
    b = P.Batch(4, P.PREC_FP32)
    for wl in wls:
        t = wl.tree
        b.add(wl.masks, t.left, t.right, t.root, t.blen, wl.Q, wl.pi, wl.cat_rates, wl.cat_weights, nni_mode=1)
    b.finalize(autotune=True)
    for _ in range(args.warmup):
        b.run()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        tot, nni = b.run()
    el = time.perf_counter() - t1
    u1, u2, ms, spi, nv = b.stats()
    return {"metric": "CLV updates/s (site*node*cat)", "value": (u1 + u2) * args.steps / el, "unit": "updates/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": f"{n_inst} GTR+G4 alignments + full NNI neighbourhoods",
                                            "updates_full": u1, "updates_nni": u2, "variants": nv,
                                            "sites_per_item": spi, "category_waves": b.launch_shape()[1],
                                            "workgroups_per_cu": b.launch_shape()[2],
                                            "device_ms": ms,
                                            "finite": bool(np.all(np.isfinite(tot)) and np.all(np.isfinite(nni)))}}


def bench_learn(args, P):
    """Learned proposal: simulate and featurise --train alignments (32 taxa x
    2000 sites, random GTR+G4 on Yule trees) on the GPU, train the
    256-512-512-12 MLP on the GPU, then fit --heldout new alignments twice on
    their true topologies (FP64 partials): Nelder-Mead from the default start
    (equal rates and frequencies, alpha 1, mean length 0.1, simplex edge 0.3)
    and from the MLP's prediction (simplex edges twice its held-out RMSE);
    both stop at an lnL spread of 1e-3 and restart with a 0.01 simplex until a
    restart gains less than that; before those restarts each coordinate is
    probed at its start value and the full-size simplex rerun from a probe
    that gains.  Each arm fits all alignments on --fit-workers
    host threads, one context each; the value is the ratio of the two arms'
    wall times."""
    from phylommand_amd import learn as L
    rng = np.random.default_rng(5)
    t0 = time.perf_counter()
    draws = L.draw_models(args.train, 32, rng)
    feats = L.simulate_features(draws, 2000, seed=11)
    t_sim = time.perf_counter() - t0
    t0 = time.perf_counter()
    prop = L.Proposal()
    losses = prop.fit(feats, L.targets(draws), schedule=[(20, 1e-3), (10, 3e-4), (10, 1e-4)])
    t_train = time.perf_counter() - t0
    print(f"[bench] learn: simulated {args.train} in {t_sim:.1f}s, trained in {t_train:.1f}s", file=sys.stderr,
          flush=True)
    held = L.draw_models(args.heldout, 32, np.random.default_rng(99))
    hf, hm = L.simulate_features(held, 2000, seed=77, want_masks=True)
    common = dict(ftol_abs=1e-3, xtol_rel=1e-4, max_evals=8000, restart_step=0.01)
    c = L.compare_fits(prop, held, hf, hm, opts=P.nm_options(**common),
                       seeded_opts=P.nm_options(**common, steps=2 * prop.x_rmse), workers=args.fit_workers)
    diff = c.lnl_seeded - c.lnl_plain
    ratio = c.wall_plain / c.wall_seeded
    return {"metric": "ML-fit wall time, plain / learned-seeded (matched lnL)", "value": ratio, "unit": "x",
            "n_gpus": 1, "steps": args.heldout, "warmup": 0,
            "ms_per_step": 1e3 * c.wall_seeded / args.heldout, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (GPU-simulated GTR+G4 alignments)",
            "config": {"workload": f"learned proposal: {args.train} training alignments, {args.heldout} held-out "
                                   f"fits, 32 taxa x 2000 sites, GTR+G4",
                       "plain_wall_s": c.wall_plain, "seeded_wall_s": c.wall_seeded, "fit_threads": c.workers,
                       "plain_fit_s_sum": float(c.sec_plain.sum()), "seeded_fit_s_sum": float(c.sec_seeded.sum()),
                       "plain_evals_mean": float(c.evals_plain.mean()),
                       "seeded_evals_mean": float(c.evals_seeded.mean()),
                       "lnl_within_0.01": int(np.sum(np.abs(diff) <= 0.01)),
                       "seeded_higher_by_more": int(np.sum(diff > 0.01)),
                       "plain_higher_by_more": int(np.sum(diff < -0.01)),
                       "max_plain_minus_seeded": float(max(0.0, -diff.min())),
                       "max_seeded_minus_plain": float(max(0.0, diff.max())),
                       "final_train_loss": float(losses[-1]), "simulate_s": t_sim, "train_s": t_train,
                       "proposal_rmse": [float(v) for v in prop.x_rmse]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 500 dna/protein, 50 otherwise)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (default 300 dna/protein: the kernel time settles after about "
                         "100 evaluations, about 20 ms of GPU work; 5 otherwise)")
    ap.add_argument("--sites", type=int, default=None, help="sites (default 100,000 DNA, 20,000 protein)")
    ap.add_argument("--taxa", type=int, default=128)
    ap.add_argument("--fp64", action="store_true", help="store partials in FP64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-protein", action="store_true", help="dna: skip the 20-state walker's extra keys")
    ap.add_argument("--workload", default="dna", choices=["dna", "protein", "nni", "batch", "learn"])
    ap.add_argument("--train", type=int, default=200_000, help="learn: training alignments")
    ap.add_argument("--heldout", type=int, default=1000, help="learn: held-out fits")
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--fit-workers", type=int, default=8,
                    help="learn: host threads fitting held-out alignments concurrently (each arm in turn)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run only a plain device-to-device copy of MIB MiB (torch), the calibration kernel "
                         "for the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--copy-time", action="store_true", help="with --copy-probe: time the copy, print GB/s")
    args = ap.parse_args()
    short = args.workload in ("dna", "protein")
    if args.steps is None:
        args.steps = 500 if short else 50
    if args.warmup is None:
        args.warmup = 300 if short else 5
    if args.copy_probe and args.copy_time:
        print(json.dumps({"copy_gbs": copy_bandwidth(args.copy_probe)}))
        return
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": args.copy_probe, "bytes_read": a.numel() * 4,
                          "bytes_written": b.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(sites shard across GPUs with one 8-byte all-reduce per evaluation; not built here)",
              file=sys.stderr)
        sys.exit(2)

    import phylommand_amd as P
    from phylommand_amd import synth

    if args.workload != "dna":
        if args.workload == "learn":
            print(json.dumps(bench_learn(args, P)))
            return
        fn = {"protein": bench_protein, "nni": bench_nni}.get(args.workload)
        out = fn(args, P, synth) if fn else bench_batch(args, P, synth, args.instances)
        print(json.dumps(out))
        return

    t0 = time.time()
    args.sites = args.sites or 100_000
    wl = synth.gtr_gamma_workload(n_tips=args.taxa, m=args.sites, mean_blen=0.05, alpha=0.5, C=4,
                                  gap_frac=0.05, seed=2)
    gen_s = time.time() - t0
    print(f"[bench] workload generated in {gen_s:.1f}s", file=sys.stderr, flush=True)
    t = wl.tree
    ctx = P.Context(0, P.PREC_FP64 if args.fp64 else P.PREC_FP32)
    ctx.set_tips(wl.masks, 4)
    ctx.set_topology(t.left, t.right, t.root)
    ctx.set_edge_lengths(t.blen)
    ctx.set_model(wl.Q, wl.pi, wl.pi, P.P_EIGEN)
    ctx.set_categories(wl.cat_rates, wl.cat_weights)
    lnl0 = ctx.loglik()
    for _ in range(args.warmup):
        ctx.loglik()
    ctx.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lnl = ctx.loglik()
        kms.append(ctx.last_kernel_ms())
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    assert np.isfinite(lnl) and lnl == lnl0, (lnl, lnl0)

    # End to end, as an optimiser calls it: new edge lengths before every
    # evaluation (eigen system and lengths copied host -> device, P rebuilt
    # on the device), then the evaluation and the total back on the host.
    blens = (t.blen, t.blen * 1.01)
    e2e_steps = max(20, min(args.steps, 200))
    seen = [None, None]
    t1 = time.perf_counter()
    for i in range(e2e_steps):
        ctx.set_edge_lengths(blens[i & 1])
        v = ctx.loglik()
        assert seen[i & 1] is None or v == seen[i & 1], (i, v, seen[i & 1])
        seen[i & 1] = v
    e2e_ms = (time.perf_counter() - t1) / e2e_steps * 1e3
    ctx.set_edge_lengths(t.blen)

    n, m, C = t.n_tips, args.sites, 4
    updates = (n - 1) * m * C
    value = updates * args.steps / elapsed
    kernel_s = float(np.mean(kms)) * 1e-3
    bpu = alg_bytes_per_update(n, 4, C) * (2 if args.fp64 else 1)
    achieved = flops_per_update(4) * updates / kernel_s / 1e12
    traffic = profiled_traffic("prune4_gtr_g4_128x100k") if (not args.fp64 and m == 100_000 and n == 128) else None
    out = {
        "metric": "CLV updates/s (site*node*cat)",
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Yule tree, GTR+G4 simulation)",
        # every site is evaluated (no pattern compression in the timed step):
        # updates are counted over sites; n_patterns is how many distinct
        # columns the alignment has
        "config": {"workload": f"GTR+G4 DNA, {n} taxa Yule, {m} sites, 5% gaps",
                   "n_taxa": n, "n_sites": m, "n_patterns": int(np.unique(wl.masks, axis=1).shape[1]),
                   "n_cat": C, "clv_storage": "f64" if args.fp64 else "f32", "arithmetic": "f64", "lnL": lnl},
        "end_to_end": {"ms_per_step": e2e_ms, "steps": e2e_steps,
                       "what": "set_edge_lengths with new lengths (host-to-device copy of the eigen system "
                               "and edge lengths, P rebuilt on the device) + loglik, alternating two length sets"},
        # What binds the DNA walker is FP64 arithmetic issue, not HBM: it keeps
        # all but a few partials in registers and LDS, so it moves a tenth of
        # the bytes a store-every-partial design would (alg_bytes_per_update),
        # and the counters show HBM far from busy (hbm_frac).  achieved is the
        # update's FP64 flops (2 P x L products + the elementwise product) per
        # second of kernel time (both launches of a tail split, gap included).
        "roofline": {"bound": "valu_fp64", "achieved": achieved, "peak": FP64_VALU_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_VALU_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "prune4_kernel<float,4,true>", "kernel_ms": kernel_s * 1e3,
                     "flops_per_update": flops_per_update(4),
                     "hbm_gbs": traffic / kernel_s / 1e9 if traffic else None,
                     "hbm_frac": traffic / kernel_s / 1e9 / HBM_PEAK_GBS if traffic else None,
                     "alg_bytes_per_update": bpu,
                     "alg_bytes_gbs": bpu * updates / kernel_s / 1e9},
    }
    try:
        out["roofline"]["copy_gbs"] = copy_bandwidth_subprocess()
    except Exception as e:  # optional calibration
        out["roofline"]["copy_gbs"] = None
        print(f"[bench] copy bandwidth not measured: {e}", file=sys.stderr)
    if not args.no_protein:
        # the 20-state walker on its own workload, timed the same way: context
        # for the DNA line, not part of its value
        try:
            r = measure_protein(P, synth, min(args.steps, 200), min(args.warmup, 300))
            out["protein_20_state"] = {"workload": "20-state protein + G4, 64 taxa Yule, 20000 sites, 3% gaps",
                                       "updates_per_s": r["value"], "ms_per_step": r["ms_per_step"],
                                       "kernel_ms": r["kernel_ms"], "lnL": r["lnL"], "roofline": r["roofline"]}
        except Exception as e:
            out["protein_20_state"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(wl)
        except Exception as e:  # the GPU figure stands on its own
            out["cpu_baseline"] = {"value": None, "error": str(e)}
        try:
            out["cpu_baseline"]["fair_cpu"] = fair_cpu(wl, out["cpu_baseline"].get("cores") or host_cores()[0])
        except Exception as e:
            out["cpu_baseline"]["fair_cpu"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
