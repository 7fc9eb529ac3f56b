"""Benchmark: relaxed networks/s for a batch of uneven networks on one MI355X.

Workload (synthetic, seeded): 512 independent Lees-Edwards triangular networks,
L uniform even in [16, 192], bond dilution p ~ U[0.55, 0.75], per-fiber bending
stiffness log-uniform in [1e-5, 1e-2], each sheared affinely to gamma ~ U[0, 0.4]
and relaxed by FIRE to max node |F| < 1e-9.  A step = relax the whole batch
from the affine start (positions restored on the device before each step).

Also measured: the plain form (one network per launch, FP64, the reference's
way of running) on a sample of the same networks, the relax kernel's algorithmic
HBM rate against MI355X's 8 TB/s, and the CPU FP64 oracle on the host cores
over a bounded sample.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FIRE = dict(alpha0=0.1, f_alpha=0.99, dt_max=0.5, f_inc=1.1, f_dec=0.5, nmin=5, mass=1.0)


def workload(n_nets, lmin, lmax, seed):
    rng = np.random.RandomState(seed)
    specs = []
    for i in range(n_nets):
        L = 2 * rng.randint(lmin // 2, lmax // 2 + 1)
        p = rng.uniform(0.55, 0.75)
        gamma = rng.uniform(0.0, 0.4)
        specs.append(dict(L=int(L), p=float(p), gamma=float(gamma), seed=1000 + i,
                          kseed=5000 + i))
    return specs


def build_networks(N, specs):
    nets = []
    for s in specs:
        n = N.Network.lattice(s["L"], s["p"], s["seed"], True).triples()
        n.fiber_kappa(1e-5, 1e-2, s["kseed"])
        nets.append(n)
    return nets


def alg_bytes(N_, E, T, per_triple_kappa=True):
    """Algorithmic bytes of one FP64 FIRE step: x, v, F read+written (96 B/node),
    clamp mask (1), bond (2 int32 + f64 rest + int8 wrap = 17), triple
    (3 int32 + 2 int8 wraps = 14) and its f64 kappa (8)."""
    return 97 * N_ + 17 * E + (22 if per_triple_kappa else 14) * T


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--networks", type=int, default=512)
    ap.add_argument("--lmin", type=int, default=16)
    ap.add_argument("--lmax", type=int, default=192)
    ap.add_argument("--fcut", type=float, default=1e-9)
    ap.add_argument("--max-steps", type=int, default=3000000)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--team", type=int, default=0)
    ap.add_argument("--tpm", type=int, default=0, help="tiles per team member")
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the batched relax shards by network "
              "across GPUs, which this build does not run (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import network_relaxation_amd as N

    specs = workload(args.networks, args.lmin, args.lmax, args.seed)
    nets = build_networks(N, specs)
    batch = N.Batch(nets, 1.0, 1.0, tile_nodes=args.tile, team=args.team,
                    tiles_per_member=args.tpm)
    prm = N.relax_params(fcut=args.fcut, max_steps=args.max_steps, criterion=N.CRIT_MAX, **FIRE)

    def one_step():
        batch.reset()
        for k, s in enumerate(specs):
            batch.shear(k, s["gamma"])
        return batch.relax(prm)

    for _ in range(args.warmup):
        one_step()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        reps = one_step()
        kernel_ms += batch.last_ms()
    wall = time.perf_counter() - t0

    conv = [r["status"] == N.CONVERGED for r in reps]
    n_conv = int(sum(conv))
    value = n_conv * args.steps / wall
    t0d = min(r["t_start"] for r in reps)
    lat = np.array([r["t_done"] - t0d for r in reps])
    steps = np.array([r["steps"] for r in reps], dtype=np.float64)

    # roofline of the relax kernel (one launch per step)
    total_bytes = 0.0
    node_steps = 0.0
    for k in range(len(specs)):
        n_, e_, t_ = batch.counts(k)
        total_bytes += reps[k]["steps"] * alg_bytes(n_, e_, t_)
        node_steps += reps[k]["steps"] * n_
    achieved = total_bytes / (kernel_ms / args.steps * 1e-3) / 1e9

    # plain form: one network per launch, FP64, on a sample
    plain = None
    if not args.no_plain and args.plain_sample > 0:
        rng = np.random.RandomState(args.seed + 1)
        idx = sorted(rng.choice(len(specs), size=min(args.plain_sample, len(specs)), replace=False))
        t_plain = 0.0
        ok = 0
        for k in idx:
            b1 = N.Batch([nets[k]], 1.0, 1.0, tiles_per_member=1)
            b1.shear(0, specs[k]["gamma"])
            ts = time.perf_counter()
            r1 = b1.relax(prm)[0]
            t_plain += time.perf_counter() - ts
            ok += r1["status"] == N.CONVERGED
            del b1
        plain = dict(networks_per_s=ok / t_plain, sample=len(idx),
                     sample_nodes=int(sum(specs[k]["L"] ** 2 for k in idx)))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(specs, reps, args)

    out = {
        "metric": "relaxed networks/s at |F|max<1e-9: batched+mixed vs plain FIRE; p99 latency",
        "value": value,
        "unit": "networks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batched: {len(specs)} Lees-Edwards triangular networks, L even "
                               f"U[{args.lmin},{args.lmax}], p U[0.55,0.75], per-fiber kappa "
                               f"log-U[1e-5,1e-2], gamma U[0,0.4] affine start, FIRE to "
                               f"|F|max<{args.fcut:g}",
                   "networks": len(specs), "nodes": int(sum(s["L"] ** 2 for s in specs)),
                   "precision": "fp64", "fire": FIRE, "parallelism": "1 GPU, persistent batch"},
        "converged": n_conv,
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "max": float(lat.max())},
        "fire_steps": {"mean": float(steps.mean()), "max": int(steps.max()),
                       "node_steps": node_steps},
        "plain_fp64": plain,
        "speedup_vs_plain": (value / plain["networks_per_s"]) if plain else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "relax_kernel", "kernel_ms": kernel_ms / args.steps,
                     "alg_bytes_per_launch": total_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def calibrate_copy(mib):
    """A plain device-to-device copy kernel of known size (torch's elementwise
    copy), so that the FETCH_SIZE / WRITE_SIZE counters can be scaled."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": 4 * n, "launches": 3}), flush=True)


def cpu_baseline(specs, reps, args):
    """FP64 oracle FIRE on the host cores over a bounded sample: a fixed number
    of FIRE steps on sampled networks (one network per thread) gives node-steps/s;
    networks/s = that rate / the batch's measured node-steps per network."""
    try:
        import oracle as O
    except Exception as e:  # the oracle is test infrastructure; report absence
        return {"value": None, "unit": "networks/s", "error": str(e)}
    cores = min(16, os.cpu_count() or 1)
    rng = np.random.RandomState(args.seed + 2)
    idx = rng.choice(len(specs), size=min(cores, len(specs)), replace=False)
    per_net_steps = 400

    def run(k):
        s = specs[k]
        o = O.Net.lattice(s["L"], s["p"], s["seed"], True).triples()
        o.fiber_kappa(1e-5, 1e-2, s["kseed"])
        pos = o.pos()
        o.shear(pos, s["gamma"])
        t = time.perf_counter()
        r = o.fire(1.0, 1.0, pos, fp=(FIRE["alpha0"], FIRE["f_alpha"], FIRE["dt_max"],
                                        FIRE["f_inc"], FIRE["f_dec"]),
                   nmin=FIRE["nmin"], mass=FIRE["mass"], fcut=args.fcut,
                   max_steps=per_net_steps, criterion=O.CRIT_MAX)
        return r["steps"] * o.n, time.perf_counter() - t

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        res = list(ex.map(run, idx))
    wall = time.perf_counter() - t0
    ns = sum(r[0] for r in res)
    rate = ns / wall  # node-steps/s over all threads
    node_steps_per_net = np.mean([reps[k]["steps"] * specs[k]["L"] ** 2 for k in range(len(specs))])
    return {"value": rate / node_steps_per_net, "unit": "networks/s", "cores": cores,
            "kind": "port",
            "sample": f"{len(idx)} sampled networks x {per_net_steps} FIRE steps on {cores} "
                      f"threads ({wall:.1f} s): {rate:.3e} node-steps/s, divided by the batch's "
                      f"measured {node_steps_per_net:.3e} node-steps per relaxed network"}


if __name__ == "__main__":
    main()
