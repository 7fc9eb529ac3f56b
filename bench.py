"""Benchmark: relaxed networks/s for a batch of uneven networks on one MI355X.

Workload (synthetic, seeded): 512 independent Lees-Edwards triangular networks,
L uniform even in [16, 192], bond dilution p ~ U[0.55, 0.75], per-fiber bending
stiffness log-uniform in [1e-5, 1e-2], each sheared affinely to gamma ~ U[0, 0.4]
and relaxed by FIRE to max node |F| < 1e-9.  A step = relax the whole batch
from the affine start (positions restored on the device before each step) in
mixed precision (FP32 stage on FP64 anchors, FP64 polish); the headline value
is that form's converged networks/s.  The untimed warmup relax plans the
device teams from estimated step counts; the timed relaxes plan them from the
step counts of the batch's last relax (as every increment of a strain ramp
after the first does); the first relax's time is reported beside the value.

Also measured, each on the same MI355X: the same batch in FP64 (and the
per-network energy difference of the mixed form from it), the plain form (one
network per launch, FP64) on a sample, a served stream (Poisson arrivals of
strain-ramp jobs through the serving loop, p50/p99 latency), the relax
kernel's algorithmic HBM rate against MI355X's 8 TB/s, and the CPU FP64 oracle
on the host cores over a bounded sample (FIRE, and beside it the reference's
small-strain linear solve, also timed on the GPU).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
T_START = time.perf_counter()


def progress(msg):
    """One progress line on stderr (the JSON result goes to stdout)."""
    print(f"[bench {time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per relax launch of the default workload (512 networks, seed 2024,
# L 16..192, |F|max < 1e-9, tuned shapes; a relax planned from the previous
# one's step counts), from the relax kernel's own request counters
# (rocprofv3 --pmc): reads 32 RDREQ_32B + 64 RDREQ_64B + 128 RDREQ_128B =
# 64 x 182,394,455 + 128 x 193,891,664,407, writes 64 WRREQ_64B + 32 (WRREQ -
# WRREQ_64B) = 64 x 102,991,165,248 + 32 x 489,643,846 (the same sums give
# exactly 1 GiB for a 1 GiB device copy).
PROFILED_TRAFFIC_BYTES = (64 * 182394455 + 128 * 193891664407 +
                          64 * 102991165248 + 32 * (103480809094 - 102991165248))
# The same counters on one L=512 network's FIRE ramp (p 0.6, kappa 1e-4):
# counted HBM bytes per FIRE step (3.4910e12 read + 1.6760e12 written over
# 131,823 steps; the tile blobs stay resident in LDS).
RAMP512_TRAFFIC_PER_STEP = (64 * 565665115 + 128 * 26990934960 +
                            64 * 25917849604 + 32 * (26457814164 - 25917849604)) / 131823
# ... and with the shape nr_autotune picks for it (512-thread workgroups on
# 512-node tiles, 2 tiles per member; profiles/relax_kernel_ramp512_fire_tuned.md)
RAMP512_TUNED_SHAPE = {"block": 512, "tile_nodes": 512, "tiles_per_member": 2}
RAMP512_TUNED_TRAFFIC_PER_STEP = (64 * 312728048 + 128 * 22721546895 +
                                  64 * 25918046213 + 32 * (26188030549 - 25918046213)) / 131824
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


def alg_bytes32(N_, E, T):
    """Algorithmic bytes of one FP32-stage step: FP64 anchor read (16 B/node),
    FP32 u, v, F read+written (48), mask (1); bond with an FP32 rest-length
    term (13 B); triple (14) and its f64 kappa (8)."""
    return 65 * N_ + 13 * E + 22 * T


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
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup (256 or 512)")
    ap.add_argument("--autotune", action="store_true",
                    help="ramp: workgroup / tile shape and tiles per member from nr_autotune")
    ap.add_argument("--team", type=int, default=0)
    ap.add_argument("--tpm", type=int, default=0, help="tiles per team member")
    ap.add_argument("--switch", type=float, default=0.0, help="mixed FP32 -> FP64 hand-over")
    ap.add_argument("--plain-sample", type=int, default=64,
                    help="networks of the batch relaxed in the plain form, one from each of "
                         "this many strata of equal size in L order (0 = skip)")
    ap.add_argument("--served", type=int, default=8, help="served requests (0 = skip)")
    ap.add_argument("--served-lmax", type=int, default=1024)
    ap.add_argument("--served-tpm", type=int, default=0,
                    help="tiles per team member of the server (0 = its default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-linear", action="store_true", help="skip the GPU linear solves")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager relaxes of the batch (default shape) instead of graph "
                         "replays of the autotuned batch's strain step")
    ap.add_argument("--no-composite", action="store_true",
                    help="skip the network + gel composite relax")
    ap.add_argument("--no-ramp", action="store_true",
                    help="skip the two-increment L=512 FIRE ramp (autotuned shape) of the "
                         "default run")
    ap.add_argument("--composite-L", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a device copy of MIB MiB after the timed work, as "
                         "the counter calibration (alone with --steps 0)")
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "single", "ramp", "large", "composite", "served"],
                    help="batched: the headline (512 uneven networks); single: one "
                         "L=32 network vs the oracle; ramp: L=512, 100 shear increments, "
                         "FIRE and CG; large: L=2048 mixed vs FP64 per increment")
    ap.add_argument("--methods", default="fire,cg",
                    help="ramp: relax methods, of fire, cg (Fletcher-Reeves), cg_pr "
                         "(Polak-Ribiere+)")
    ap.add_argument("--increments", type=int, default=0,
                    help="ramp/large: increments to run (0 = 20 / 2)")
    ap.add_argument("--precisions", default="mixed,fp64",
                    help="large: precisions to run, of mixed and fp64 (one alone writes its "
                         "per-increment reports for a comparison made elsewhere)")
    args = ap.parse_args()
    if args.calibrate_copy and args.workload == "batched" and args.steps == 0:
        print(json.dumps(calibrate_copy(args.calibrate_copy)), flush=True)
        return
    if args.workload != "batched":
        if args.gpus != 1:
            print("bench.py runs on one GPU only (use --gpus 1)", file=sys.stderr)
            sys.exit(2)
        import network_relaxation_amd as N
        leg = {"single": single_leg, "ramp": ramp_leg, "large": large_leg,
               "composite": composite_leg, "served": served_leg}[args.workload]
        res = leg(N, args)
        if args.calibrate_copy:  # after the timed work, for the counter passes
            res["copy_calibration"] = calibrate_copy(args.calibrate_copy)
        print(json.dumps(res), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the batched relax shards by network "
              "across GPUs, which this build does not run (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import network_relaxation_amd as N

    specs = workload(args.networks, args.lmin, args.lmax, args.seed)
    nets = build_networks(N, specs)
    progress(f"built {len(nets)} networks")
    prm = N.relax_params(fcut=args.fcut, max_steps=args.max_steps, criterion=N.CRIT_MAX, **FIRE)
    shape = dict(tile_nodes=args.tile, team=args.team, tiles_per_member=args.tpm, block=args.block)

    cold_ms = []  # kernel time of the first relax (teams planned from estimated step counts)

    def run_batch(batch, steps, warmup):
        def one_step():
            batch.reset()
            for k, s in enumerate(specs):
                batch.shear(k, s["gamma"])
            return batch.relax(prm)
        cold_ms.clear()
        for w in range(warmup):
            one_step()
            cold_ms.append(batch.last_ms())
            progress(f"warmup step {w + 1}/{warmup}: {batch.last_ms():.0f} ms")
        kernel_ms = 0.0
        t0 = time.perf_counter()
        for k in range(steps):
            reps = one_step()
            kernel_ms += batch.last_ms()
            progress(f"timed step {k + 1}/{steps}: {batch.last_ms():.0f} ms")
        return reps, time.perf_counter() - t0, kernel_ms / steps

    xd = np.array([s["gamma"] * n.info().height for s, n in zip(specs, nets)])

    def run_graph(batch, steps, warmup):
        # eager warm-up relaxes (the first planned from estimated step counts,
        # the others from the last relax's), then the strain step captured and
        # replayed; the capture keeps the team plan of the last relax, so at
        # least two relaxes precede it; the first replay must equal the last
        # eager relax bit for bit
        cold_ms.clear()
        eager = None
        warmup = max(2, warmup)
        for w in range(warmup):
            batch.reset()
            batch.strain_all(xd)
            eager = batch.relax(prm)
            cold_ms.append(batch.last_ms())
            progress(f"warmup step {w + 1}/{warmup}: {batch.last_ms():.0f} ms")
        batch.graph_capture(xd, 0.0, prm)
        kernel_ms = 0.0
        same = None
        t0 = time.perf_counter()
        for k in range(steps):
            batch.reset()
            reps = batch.graph_replay()
            kernel_ms += batch.last_ms()
            if same is None:
                same = all((a["steps"], a["es"], a["eb"]) == (b["steps"], b["es"], b["eb"])
                           for a, b in zip(eager, reps))
            progress(f"timed step {k + 1}/{steps} (graph replay): {batch.last_ms():.0f} ms")
        return reps, time.perf_counter() - t0, kernel_ms / steps, same

    # headline: batched + mixed precision, as config 4 states it: shapes from
    # nr_autotune (unless given) and the timed strain steps replayed from a
    # captured HIP graph (each one from the unstrained state, after the
    # untimed eager warm-up relaxes); --no-graph times eager relaxes instead
    use_graph = not args.no_graph
    tuned_shape, tune_info = shape, None
    if use_graph and not (args.tile or args.team or args.tpm or args.block):
        t0 = time.perf_counter()
        tuned, predicted = N.autotune(nets, 1.0, 1.0, precision=N.PREC_MIXED, probe_steps=400)
        tune_info = {"shape": tuned, "predicted_ms": predicted,
                     "seconds": time.perf_counter() - t0}
        tuned_shape = dict(shape, **tuned)
        progress(f"autotune chose {tuned} in {tune_info['seconds']:.1f} s")
    mixed = N.Batch(nets, 1.0, 1.0, precision=N.PREC_MIXED, mixed_switch=args.switch,
                    **tuned_shape)
    progress("batched mixed: created")
    if use_graph:
        reps, wall, kernel_ms, replay_same = run_graph(mixed, args.steps, args.warmup)
    else:
        reps, wall, kernel_ms = run_batch(mixed, args.steps, args.warmup)
        replay_same = None
    n_conv = int(sum(r["status"] == N.CONVERGED and r["fmax"] < args.fcut for r in reps))
    value = n_conv * args.steps / wall
    t0d = min(r["t_start"] for r in reps)
    lat = np.array([r["t_done"] - t0d for r in reps])
    steps = np.array([r["steps"] for r in reps], dtype=np.float64)
    total_bytes = 0.0
    node_steps = 0.0
    for k in range(len(specs)):
        n_, e_, t_ = mixed.counts(k)
        s64 = reps[k]["fp64_steps"]
        s32 = reps[k]["steps"] - s64
        total_bytes += s32 * alg_bytes32(n_, e_, t_) + s64 * alg_bytes(n_, e_, t_)
        node_steps += reps[k]["steps"] * n_
    achieved = total_bytes / (kernel_ms * 1e-3) / 1e9
    profiled = (args.networks, args.lmin, args.lmax, args.seed, args.fcut, args.tile, args.team,
                args.tpm, args.switch, args.block) == (512, 16, 192, 2024, 1e-9, 0, 0, 0, 0.0, 0)
    traffic = PROFILED_TRAFFIC_BYTES if profiled else None
    e_mixed = np.array([r["es"] + r["eb"] for r in reps])
    first_ms = cold_ms[0] if cold_ms else None
    del mixed

    graph = None if not use_graph else {
        "timed_steps": "HIP graph replays of the captured strain step (nr_batch_graph_capture, "
                       "nr_batch_graph_replay), each from the unstrained state",
        "autotune": tune_info, "shape": tuned_shape,
        "eager_warmup_relaxes": max(2, args.warmup),
        "first_replay_equals_last_eager_relax": replay_same}

    # the same batch in FP64 (timed like the mixed form: an untimed first
    # relax, then a relax planned from its step counts), and the mixed form's
    # energy error against it
    fp64 = None
    if not args.no_fp64:
        b64 = N.Batch(nets, 1.0, 1.0, precision=N.PREC_FP64, **shape)
        progress("batched fp64: created")
        r64, wall64, kms64 = run_batch(b64, 1, 1)
        cold64 = cold_ms[0]
        n64 = sum(r["status"] == N.CONVERGED for r in r64)
        e64 = np.array([r["es"] + r["eb"] for r in r64])
        rel = np.abs(e_mixed - e64) / np.abs(e64)
        fp64 = {"networks_per_s": n64 / wall64,
                "kernel_ms": kms64,
                "first_relax": {"kernel_ms": cold64, "networks_per_s": n64 / (cold64 * 1e-3)},
                "node_steps_per_s": sum(r["steps"] * s["L"] ** 2 for r, s in zip(r64, specs))
                                    / (kms64 * 1e-3),
                "mixed_energy_rel_diff": {"median": float(np.median(rel)),
                                          "p90": float(np.percentile(rel, 90)),
                                          "max": float(rel.max()),
                                          "n_over_1e-10": int((rel > 1e-10).sum()),
                                          "n": len(rel)}}
        del b64

    # plain form: one network per launch, FP64, on a stratified sample: the
    # batch sorted by L (then p) is cut into plain_sample strata of equal
    # size and one network is drawn from each, so the sample follows the
    # batch's size distribution; each sampled time stands for its stratum
    # (an unbiased estimate of the time to relax the whole batch this way)
    plain = None
    if not args.no_plain and args.plain_sample > 0:
        rng = np.random.RandomState(args.seed + 1)
        order = sorted(range(len(specs)), key=lambda k: (specs[k]["L"], specs[k]["p"]))
        strata = np.array_split(np.array(order), min(args.plain_sample, len(specs)))
        idx = [int(rng.choice(st)) for st in strata]
        t_est = 0.0
        t_sample = 0.0
        ok = 0
        plain_node_steps = 0.0
        for k, st in zip(idx, strata):
            b1 = N.Batch([nets[k]], 1.0, 1.0, tiles_per_member=1)
            b1.shear(0, specs[k]["gamma"])
            ts = time.perf_counter()
            r1 = b1.relax(prm)[0]
            dt = time.perf_counter() - ts
            t_sample += dt
            t_est += dt * len(st)
            plain_node_steps += r1["steps"] * specs[k]["L"] ** 2
            ok += r1["status"] == N.CONVERGED
            del b1
        progress(f"plain: {len(idx)} stratified networks in {t_sample:.1f} s")
        plain = dict(networks_per_s=len(specs) / t_est * ok / len(idx), sample=len(idx),
                     sample_wall_s=t_sample, converged=int(ok),
                     node_steps_per_s=plain_node_steps / t_sample,
                     sample_mean_nodes=float(np.mean([specs[k]["L"] ** 2 for k in idx])),
                     batch_mean_nodes=float(np.mean([s["L"] ** 2 for s in specs])),
                     estimate="one network from each of the strata of equal size of the "
                              "batch in L order; networks/s = batch size / (sum of sampled "
                              "times x stratum sizes)",
                     form="one network per launch, FP64, widest team (tiles_per_member=1)")

    served = None
    if args.served > 0:
        served = served_leg(N, args)

    lin = None if args.no_linear else gpu_linear(N, specs, linear_sample(specs, args))
    comp = None if args.no_composite else composite_leg(N, args)
    # the single large network (L = 512), its first two shear increments with
    # the shape nr_autotune picks for it
    ramp = None if args.no_ramp else ramp_leg(
        N, argparse.Namespace(**dict(vars(args), increments=2, methods="fire", autotune=True)))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(specs, reps, args)
        progress("cpu baseline done")

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
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"batched: {len(specs)} Lees-Edwards triangular networks, L even "
                               f"U[{args.lmin},{args.lmax}], p U[0.55,0.75], per-fiber kappa "
                               f"log-U[1e-5,1e-2], gamma U[0,0.4] affine start, FIRE to "
                               f"|F|max<{args.fcut:g}, mixed precision (FP32 stage on FP64 "
                               f"anchors, FP64 polish)",
                   "networks": len(specs), "nodes": int(sum(s["L"] ** 2 for s in specs)),
                   "precision": "mixed", "fire": FIRE, "parallelism": "1 GPU, persistent batch",
                   "timed_steps": "graph replays, autotuned shape" if use_graph else "eager"},
        "converged": n_conv,
        "first_relax": None if first_ms is None else {
            "kernel_ms": first_ms, "networks_per_s": n_conv / (first_ms * 1e-3),
            "note": "the untimed first relax of the batch, teams planned from estimated step "
                    "counts; timed relaxes are planned from the step counts of the last relax"},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "max": float(lat.max())},
        "fire_steps": {"mean": float(steps.mean()), "max": int(steps.max()),
                       "fp64_mean": float(np.mean([r["fp64_steps"] for r in reps])),
                       "node_steps": node_steps},
        "node_steps_per_s": node_steps / (kernel_ms * 1e-3),
        "graph_replay": graph,
        "batched_fp64": fp64,
        "plain_fp64": plain,
        "speedup_vs_plain": (value / plain["networks_per_s"]) if plain else None,
        "speedup_vs_plain_first_relax": (n_conv / (first_ms * 1e-3) / plain["networks_per_s"])
                                        if plain and first_ms else None,
        "served": served,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "relax_kernel<256,1>", "kernel_ms": kernel_ms,
                     "alg_bytes_per_launch": total_bytes},
        "cpu_baseline": cpu,
        "linear_response_gpu": lin,
        "composite": comp,
        "ramp_l512": ramp,
    }
    if args.calibrate_copy:  # after the timed work, for the counter passes
        out["copy_calibration"] = calibrate_copy(args.calibrate_copy)
    print(json.dumps(out), flush=True)


def single_leg(N, args):
    """One L=32 Lees-Edwards network (p 0.65, kappa 1e-3, affine shear 0.01)
    relaxed by FIRE to |F|max < 1e-10 (dt_max 0.1), against the CPU oracle."""
    import oracle as O
    fire = dict(FIRE, dt_max=0.1)
    n = N.Network.lattice(32, 0.65, 1, True).triples()
    b = N.Batch([n], 1.0, 1e-3, tiles_per_member=1)
    prm = N.relax_params(fcut=1e-10, criterion=N.CRIT_MAX, **fire)
    times = []
    for _ in range(3):
        b.reset()
        b.shear(0, 0.01)
        t = time.perf_counter()
        r = b.relax(prm)[0]
        times.append(time.perf_counter() - t)
    o = O.Net.lattice(32, 0.65, 1, True).triples()
    pos = o.pos()
    o.shear(pos, 0.01)
    t = time.perf_counter()
    ro = o.fire(1.0, 1e-3, pos, fp=(fire["alpha0"], fire["f_alpha"], fire["dt_max"],
                                    fire["f_inc"], fire["f_dec"]), fcut=1e-10)
    t_cpu = time.perf_counter() - t
    e, eo = r["es"] + r["eb"], ro["es"] + ro["eb"]
    return {"metric": "config 1: time to relax one L=32 network, FP64 FIRE, |F|max<1e-10",
            "value": min(times) * 1e3, "unit": "ms", "higher_is_better": False,
            "steps": r["steps"], "oracle_steps": ro["steps"], "fmax": r["fmax"],
            "energy_rel_vs_oracle": abs(e - eo) / eo, "kernel_ms": b.last_ms(),
            "cpu_oracle_ms": t_cpu * 1e3, "cpu_cores": 1}


def ramp_leg(N, args):
    """L=512 (p 0.6, kappa 1e-4) Lees-Edwards network, shear ramp in
    increments of 0.005 (100 reach 0.5; 20 by default), each relaxed to
    |F|max < fcut from the previous state, by each of --methods (FIRE, CG
    Fletcher-Reeves, CG Polak-Ribiere+), five increments per device launch
    (nr_batch_ramp)."""
    ninc = args.increments or 20
    n = N.Network.lattice(512, 0.6, 2, True).triples()
    nodes, edges, tris = n.info().n_nodes, n.info().n_bonds, n.info().n_triples
    incs = np.array([(0.005 * n.info().height, 0.0)] * ninc)
    out = {"metric": f"config 2: node-updates/s, L=512 shear ramp of {ninc} x 0.005",
           "unit": "node-updates/s", "higher_is_better": True,
           "nodes": nodes, "bonds": edges, "triples": tris}
    methods = {"fire": N.METHOD_FIRE, "cg": N.METHOD_CG, "cg_pr": N.METHOD_CG_PR}
    shape = dict(block=args.block, tile_nodes=args.tile, team=args.team,
                 tiles_per_member=args.tpm)
    if args.autotune:
        t0 = time.perf_counter()
        tuned, predicted = N.autotune([n], 1.0, 1e-4, precision=N.PREC_FP64, probe_steps=2000,
                                      team=args.team)
        shape.update(tuned)
        out["autotune"] = {"shape": tuned, "predicted_ms": predicted,
                           "seconds": time.perf_counter() - t0, "probe_steps": 2000}
        progress(f"ramp: autotune chose {tuned} in {out['autotune']['seconds']:.1f} s")
    # counted bytes per FIRE step, for the two shapes whose counters were read
    explicit = args.block or args.tile or args.team or args.tpm
    ramp_traffic = None
    if not explicit and not args.autotune:
        ramp_traffic = RAMP512_TRAFFIC_PER_STEP
    elif not explicit and out["autotune"]["shape"] == RAMP512_TUNED_SHAPE:
        ramp_traffic = RAMP512_TUNED_TRAFFIC_PER_STEP
    for name in args.methods.split(","):
        method = methods[name]
        b = N.Batch([n], 1.0, 1e-4, **shape)
        prm = N.relax_params(fcut=args.fcut, criterion=N.CRIT_MAX, method=method,
                             max_steps=1000000, **FIRE)
        t = time.perf_counter()
        reps, kms = [], 0.0
        for c0 in range(0, ninc, 5):  # five increments per launch, a progress line each
            reps += b.ramp(incs[c0:c0 + 5], prm)[0]
            kms += b.last_ms()
            progress(f"ramp {name}: {len(reps)}/{ninc} increments, "
                     f"{sum(r['steps'] for r in reps)} steps")
        wall = time.perf_counter() - t
        steps = sum(r["steps"] for r in reps)
        n_, e_, t_ = b.counts(0)
        byt = steps * alg_bytes(n_, e_, t_, per_triple_kappa=False)
        achieved = byt / (kms * 1e-3) / 1e9
        out[name] = {"roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                                  "traffic": ramp_traffic * steps
                                  if name == "fire" and ramp_traffic else None,
                                  "note": "traffic: counted bytes per FIRE step of this "
                                          "network x steps"},
                     "node_updates_per_s": nodes * steps / (kms * 1e-3),
                     "steps_per_increment": [r["steps"] for r in reps],
                     "steps": steps, "kernel_ms": kms, "wall_s": wall,
                     "converged": sum(r["status"] == N.CONVERGED for r in reps),
                     "hbm_alg_GBs": byt / (kms * 1e-3) / 1e9,
                     "hbm_frac": byt / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "final_energy": reps[-1]["es"] + reps[-1]["eb"],
                     "final_shear_stress": reps[-1]["sxy"]}
        del b
    first = args.methods.split(",")[0]
    out["value"] = out[first]["node_updates_per_s"]
    if "fire" in out and "cg" in out:
        e1, e2 = out["fire"]["final_energy"], out["cg"]["final_energy"]
        out["fire_vs_cg_final_energy_rel"] = abs(e1 - e2) / e1
    return out


def large_leg(N, args):
    """L=2048 (p 0.62, per-fiber kappa log-U[1e-5,1e-2]) Lees-Edwards network,
    shear increments of 0.005 from the unstrained state, FIRE to |F|max < 1e-9,
    mixed precision (FP32 stage + FP64 polish) against FP64, per increment:
    energy and shear stress relative error, FIRE steps.  Run through the
    serving loop (20,000-step launches, bit-identical to one launch) so that
    the minutes-long relaxes report progress."""
    ninc = args.increments or 2
    n = N.Network.lattice(2048, 0.62, 3, True).triples().fiber_kappa(1e-5, 1e-2, 3)
    progress("large: network built")
    incs = np.array([(0.005 * n.info().height, 0.0)] * ninc)
    prm = N.relax_params(fcut=args.fcut, criterion=N.CRIT_MAX, max_steps=args.max_steps,
                         **FIRE)
    res = {}
    precs = [x for x in (("mixed", N.PREC_MIXED), ("fp64", N.PREC_FP64))
             if x[0] in args.precisions.split(",")]
    for name, prec in precs:
        srv = N.Server(precision=prec, slice_steps=20000, hbm_bytes=4 << 30, block=args.block,
                       tile_nodes=args.tile, team=args.team)
        t0 = time.perf_counter()
        srv.submit(n, 1.0, 1.0, prm, incs)
        out = []
        last = t0
        while not out:
            srv.step()
            out = srv.results()
            if time.perf_counter() - last > 10.0:
                last = time.perf_counter()
                progress(f"large {name}: {srv.stats()['slices']} launches")
        info = srv.stats()
        res[name] = (out[0]["reports"], info["busy_ms"], time.perf_counter() - t0)
        progress(f"large {name}: done in {res[name][2]:.1f} s")
        del srv
    if len(res) == 1:
        name, (reps, busy_ms, wall) = next(iter(res.items()))
        return {"metric": f"config 3: L=2048 {name}, {ninc} shear increments of 0.005",
                "unit": "s", "value": busy_ms * 1e-3, "higher_is_better": False,
                "nodes": n.info().n_nodes,
                "increments": [{"gamma": 0.005 * (i + 1), "energy": r["es"] + r["eb"],
                                "sxy": r["sxy"], "fmax": r["fmax"], "steps": r["steps"],
                                "fp64_steps": r["fp64_steps"], "status": r["status"]}
                               for i, r in enumerate(reps)]}
    rm, rf = res["mixed"][0], res["fp64"][0]
    rows = []
    for i in range(min(len(rm), len(rf))):
        em, ef = rm[i]["es"] + rm[i]["eb"], rf[i]["es"] + rf[i]["eb"]
        rows.append({"gamma": 0.005 * (i + 1), "energy_rel": abs(em - ef) / ef,
                     "stress_rel": abs(rm[i]["sxy"] - rf[i]["sxy"]) / abs(rf[i]["sxy"]),
                     "steps_mixed": rm[i]["steps"], "fp64_polish_steps": rm[i]["fp64_steps"],
                     "steps_fp64": rf[i]["steps"], "status": [rm[i]["status"], rf[i]["status"]]})
    return {"metric": f"config 3: L=2048 mixed vs FP64, {ninc} shear increments of 0.005",
            "unit": "s", "value": res["mixed"][1] * 1e-3, "higher_is_better": False,
            "fp64_s": res["fp64"][1] * 1e-3, "nodes": n.info().n_nodes,
            "max_energy_rel": max(r["energy_rel"] for r in rows),
            "max_stress_rel": max(r["stress_rel"] for r in rows), "increments": rows}


def served_leg(N, args):
    """The served stream through the serving loop: requests with L
    log-uniform in [16, served_lmax] (1024 by default), p U[0.55,0.75],
    per-fiber kappa, 1-20 shear increments of U[0.01,0.05], mixed precision.
    First all requests at once (saturated throughput X), then Poisson
    arrivals at 0.7 X with latency = arrival -> last increment done; the
    mixed results of every 4th job against an FP64 ramp of it."""
    rng = np.random.RandomState(args.seed + 5)
    n = args.served
    prm = N.relax_params(fcut=args.fcut, max_steps=args.max_steps, criterion=N.CRIT_MAX, **FIRE)
    reqs = []
    for i in range(n):
        L = int(round(np.exp(rng.uniform(np.log(16), np.log(args.served_lmax))) / 2)) * 2
        L = max(L, 16)
        p = rng.uniform(0.55, 0.75)
        ninc = int(rng.randint(1, 21))
        d = rng.uniform(0.01, 0.05, size=ninc)
        net = N.Network.lattice(L, p, 9000 + i, True).triples().fiber_kappa(1e-5, 1e-2, 7000 + i)
        yr = net.info().height
        incs = np.stack([d * yr, np.zeros(ninc)], axis=1)
        reqs.append((net, incs))
    progress(f"served: built {n} requests")

    def run(arrivals):
        srv = N.Server(precision=N.PREC_MIXED, slice_steps=1000,
                       tiles_per_member=args.served_tpm, block=args.block, tile_nodes=args.tile)
        t0 = srv.clock()
        i = done = 0
        res = {}
        ids = {}
        last = time.perf_counter()
        while done < n:
            if time.perf_counter() - last > 5.0:
                last = time.perf_counter()
                info = srv.stats()
                progress(f"served: submitted {i}/{n} done {done} active {info['active']} "
                         f"waiting {info['waiting']} launches {info['slices']}")
            now = srv.clock() - t0
            while i < n and arrivals[i] <= now:
                ids[srv.submit(reqs[i][0], 1.0, 1.0, prm, reqs[i][1])] = i
                i += 1
            act, wait = srv.step()
            for r in srv.results():
                res[ids[r["job"]]] = r
                done += 1
            if act == 0 and wait == 0 and i < n:
                time.sleep(max(0.0, arrivals[i] - (srv.clock() - t0)))
        wall = max(r["t_done"] for r in res.values()) - t0
        lat = np.array([res[k]["t_done"] - res[k]["t_submit"] for k in range(n)])
        ok = sum(res[k]["status"] == N.CONVERGED for k in range(n))
        info = srv.stats()
        return res, wall, lat, ok, info

    _, wall_sat, _, ok_sat, _ = run(np.zeros(n))
    x_sat = ok_sat / wall_sat
    progress(f"served: saturated {x_sat:.2f} jobs/s")
    lam = 0.7 * x_sat
    arr = np.cumsum(rng.exponential(1.0 / lam, size=n))
    res, wall, lat, ok, info = run(arr)
    progress(f"served: poisson at {lam:.2f}/s done")
    # mixed vs FP64, per increment, on a sample of jobs (FP64 ramp on the GPU)
    errs = []
    for k in range(0, n, max(1, n // 4)):
        b = N.Batch([reqs[k][0]], 1.0, 1.0, precision=N.PREC_FP64)
        ramp = b.ramp(reqs[k][1], prm)[0]
        e_m = np.array([x["es"] + x["eb"] for x in res[k]["reports"]])
        e_f = np.array([x["es"] + x["eb"] for x in ramp])
        errs.extend(list(np.abs(e_m - e_f) / np.abs(e_f)))
        del b
    errs = np.array(errs)
    out = {"requests": n, "converged": int(ok),
            "workload": f"L log-U[16,{args.served_lmax}], p U[0.55,0.75], kappa log-U[1e-5,1e-2] "
                        f"per fiber, 1-20 shear increments U[0.01,0.05], mixed precision, "
                        f"1000-step slices",
            "increments": int(sum(len(r[1]) for r in reqs)),
            "saturated_networks_per_s": x_sat,
            "poisson_rate_per_s": lam,
            "networks_per_s": ok / wall,
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(lat.max())},
            "launches": info["slices"], "device_busy_frac": info["busy_ms"] * 1e-3 / wall,
            "mixed_vs_fp64_energy_rel": {"median": float(np.median(errs)), "max": float(errs.max()),
                                         "n_over_1e-10": int((errs > 1e-10).sum()),
                                         "n": int(len(errs))}}
    if args.workload == "served":
        out.update({"metric": "served relax jobs/s (Poisson arrivals at 0.7 x saturated) and "
                              "p50/p99 latency", "value": out["networks_per_s"],
                    "unit": "jobs/s", "n_gpus": 1, "higher_is_better": True,
                    "dtype": "f32+f64", "data": "synthetic"})
    return out


# HBM bytes per composite FIRE step at L = 512 from the kernels' request
# counters (32 RDREQ_32B + 64 RDREQ_64B + 128 RDREQ_128B read, 64 WRREQ_64B +
# 32 (WRREQ - WRREQ_64B) written; profiles/mesh_kernels_composite512.md):
# mesh_tile, then mesh_fire.
COMPOSITE512_TRAFFIC_BYTES = ((64 * 2258 + 128 * 678279 + 64 * 221311 + 32 * (223848 - 221311)) +
                              (64 * 423 + 128 * 97559 + 64 * 130161 + 32 * (130162 - 130161)))


def composite_bytes(n, e, t, f):
    """Algorithmic bytes of one composite FIRE step (two launches): the tile
    pass reads each element's record once (bond: 2 point indices, rest length
    and x image offset, 24 B; triple: 3 indices and 2 offsets, 28 B; facet: 3
    indices, the inverse rest matrix and the area, 68 B), each point's position
    (16 B), velocity (16 B) and clamp code (1 B) and writes its F and v (32 B);
    mesh_fire reads v, F, x and writes v, x (80 B per point)."""
    return 24 * e + 28 * t + 68 * f + n * (16 + 16 + 1 + 32 + 80)


def composite_leg(N, args):
    """The network + FEM-gel composite (network_mesh_relax.cu): one synthetic
    composite of side L (ks 1, kb 1e-3, bmod 0.5, smod 0.25, dilation 0.05)
    sheared by 0.05 from its dilated start and relaxed by FIRE to
    |F|_rms < 1e-9 (the reference's criterion), FP64.  A first relax (graph
    capture included) is untimed; the second, from the same start, is timed
    with HIP events.  CPU: the oracle's composite FIRE on one thread for a
    bounded number of steps."""
    L = args.composite_L
    ks, kb, bmod, smod, dil, gamma = 1.0, 1e-3, 0.5, 0.25, 0.05, 0.05
    mesh = N.Mesh.lattice(L, 0.6, args.seed)
    c = N.Composite(mesh, ks, kb, bmod, smod, dil)
    info = c.info()
    start = c.positions()
    prm = N.relax_params(fcut=1e-9, criterion=N.CRIT_RMS, max_steps=2000000, **FIRE)
    runs = []
    for _ in range(2):
        c.set_positions(start)
        c.affine(info.range, gamma * info.range, 0.0)
        runs.append(c.relax(prm))
    r = runs[-1]
    n, e, t, f = info.n_points, info.n_bonds, info.n_triples, info.n_facets
    step_bytes = composite_bytes(n, e, t, f)
    achieved = step_bytes * r["steps"] / (r["ms"] * 1e-3) / 1e9
    progress(f"composite L={L}: {r['steps']} steps in {r['ms']:.1f} ms")
    out = {"workload": f"composite L={L}: network p 0.6 on a gel mesh, ks {ks}, kb {kb}, bmod "
                       f"{bmod}, smod {smod}, dilation {dil}, shear {gamma}, FIRE to "
                       f"|F|_rms<1e-9, FP64",
           "points": n, "bonds": e, "triples": t, "facets": f, "pref_d": info.pref_d,
           "steps": r["steps"], "status": r["status"], "device_ms": r["ms"],
           "first_relax_ms": runs[0]["ms"], "composites_per_s": 1e3 / r["ms"],
           "node_steps_per_s": n * r["steps"] / (r["ms"] * 1e-3),
           "us_per_step": r["ms"] * 1e3 / max(r["steps"], 1),
           "energies": {"stretch": r["es"], "bend": r["eb"], "gel": r["eg"]},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": COMPOSITE512_TRAFFIC_BYTES if L == 512 else None,
                        "alg_bytes_per_step": step_bytes,
                        "note": "two launches per FIRE step (fused element + gather tile "
                                "pass, FIRE update); achieved = algorithmic bytes of the step "
                                "x steps / device time; traffic = counted HBM bytes per step"}}
    if not getattr(args, "no_cpu", False):
        try:
            import oracle as O
            import tempfile
            with tempfile.TemporaryDirectory() as d:
                path = os.path.join(d, "c.nmc")
                mesh.write_nmc(path)
                o = O.Mesh.read(path)
                pos, pd = o.setup(ks, kb, bmod, dil)
                o.affine(pos, gamma * info.range, 0.0, info.range)
                cap = 500
                t0 = time.perf_counter()
                ro = o.fire(ks, kb, smod, bmod, pd, pos, fp=(FIRE["alpha0"], FIRE["f_alpha"],
                            FIRE["dt_max"], FIRE["f_inc"], FIRE["f_dec"]), nmin=FIRE["nmin"],
                            mass=FIRE["mass"], fcut=1e-9, max_steps=cap)
                wall = time.perf_counter() - t0
            rate = ro["steps"] * n / wall
            out["cpu_baseline"] = {
                "value": rate / (n * r["steps"]), "unit": "composites/s", "cores": 1,
                "kind": "port", "node_steps_per_s": rate,
                "sample": f"oracle composite FIRE, {ro['steps']} steps on one thread "
                          f"({wall:.1f} s), divided by the device relax's {r['steps']} steps"}
            progress(f"composite cpu sample done ({wall:.1f} s)")
        except Exception as ex:  # the oracle is test infrastructure; report absence
            out["cpu_baseline"] = {"value": None, "error": str(ex)}
    if args.workload == "composite":
        out.update({"metric": "relaxed composites/s (network + gel, |F|_rms<1e-9)",
                    "value": out["composites_per_s"], "unit": "composites/s", "n_gpus": 1,
                    "higher_is_better": True, "dtype": "f64", "data": "synthetic"})
    return out


def calibrate_copy(mib):
    """A plain device-to-device copy of known size (the HIP runtime's copy
    kernel, three launches), so that the FETCH_SIZE / WRITE_SIZE counters of
    the relax kernel can be scaled by what they read for it."""
    import ctypes as C
    path = "/opt/rocm/lib/libamdhip64.so"
    hip = C.CDLL(path if os.path.exists(path) else "libamdhip64.so")
    n = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for rc in (hip.hipMalloc(C.byref(a), n), hip.hipMalloc(C.byref(b), n),
               hip.hipMemset(a, 1, n)):
        if rc:
            raise RuntimeError(f"HIP error {rc} in the copy calibration")
    for _ in range(3):
        if hip.hipMemcpy(b, a, n, 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed in the copy calibration")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    progress(f"copy calibration: 3 x {mib} MiB device-to-device")
    return {"calibration_copy_bytes": mib << 20, "launches": 3}


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
    per_net_steps = 6000

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
    progress(f"cpu baseline: FIRE sample done ({wall:.1f} s)")
    return {"value": rate / node_steps_per_net, "unit": "networks/s", "cores": cores,
            "kind": "port", "host": host_cpu(),
            "node_steps_per_s_per_thread": rate / cores,
            "sample": f"{len(idx)} sampled networks x {per_net_steps} FIRE steps on {cores} "
                      f"threads ({wall:.1f} s): {rate:.3e} node-steps/s, divided by the batch's "
                      f"measured {node_steps_per_net:.3e} node-steps per relaxed network",
            "linear": cpu_linear(specs, linear_sample(specs, args), cores)}


def host_cpu():
    """The host's logical CPU count and CPU model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def linear_sample(specs, args):
    """Four networks of the batch for the small-strain linear solves (the
    same ones on the CPU and on the GPU)."""
    rng = np.random.RandomState(args.seed + 3)
    return sorted(rng.choice(len(specs), size=min(4, len(specs)), replace=False))


def gpu_linear(N, specs, idx):
    """The small-strain linear solve (CG on the device, stiffness assembled
    on the host) of the linear CPU baseline's networks at gamma = 0.01: one
    at a time (nr_linear_response), then all in one device batch
    (nr_linear_response_batch, as nrelax_linear solves a directory's load
    steps; each result equals the single solve's bit for bit)."""
    nets = []
    for k in idx:
        s = specs[k]
        n = N.Network.lattice(s["L"], s["p"], s["seed"], True).triples()
        n.fiber_kappa(1e-5, 1e-2, s["kseed"])
        nets.append((s["L"], n))

    def one(item):
        L, n = item
        t = time.perf_counter()
        r = N.linear_response(n, 1.0, 1.0, 0.01, max_iter=20000)
        return dict(L=L, seconds=time.perf_counter() - t, device_ms=r["device_ms"],
                    cg_iters=r["iters"], residual=r["residual"], energy=r["es"] + r["eb"])

    out = [one(item) for item in nets]
    jobs = [dict(net=n, ks=1.0, kb=1.0, shear=0.01, max_iter=20000) for _, n in nets]
    t0 = time.perf_counter()
    bat = N.linear_response_batch(jobs)
    wall = time.perf_counter() - t0
    same = all(o["cg_iters"] == r["iters"] and o["energy"] == r["es"] + r["eb"]
               for o, r in zip(out, bat))
    progress(f"linear solves on the GPU done ({wall:.2f} s as one batch)")
    return {"solves_per_s": len(bat) / wall,
            "sequential_solves_per_s": len(out) / sum(o["seconds"] for o in out),
            "gamma": 0.01, "cg_iter_cap": 20000, "networks": out,
            "batch": {"wall_s": wall, "device_ms": bat[0]["device_ms"] if bat else 0.0,
                      "equals_single_solves": same},
            "note": "host assembly + device CG, seconds include the assembly; "
                    "solves_per_s: the sample's networks in one device batch; "
                    "sequential_solves_per_s: one call per network"}


def cpu_linear(specs, idx, cores):
    """The reference's other CPU path, the small-strain linear solve
    (sparse_relax.cpp), restated in the oracle: stiffness assembly and a CG
    solve of the reduced system at gamma = 0.01, one network at a time on all
    threads.  A different problem from FIRE at finite strain (linear response),
    reported beside it, not as the baseline value."""
    import oracle as O
    out = []
    t0 = time.perf_counter()
    for k in idx:
        s = specs[k]
        o = O.Net.lattice(s["L"], s["p"], s["seed"], True).triples()
        o.fiber_kappa(1e-5, 1e-2, s["kseed"])
        t = time.perf_counter()
        r = o.linear_response(1.0, 1.0, 0.01, max_iter=20000)
        out.append(dict(L=s["L"], seconds=time.perf_counter() - t, cg_iters=r["iters"],
                        residual=r["residual"]))
    wall = time.perf_counter() - t0
    progress(f"cpu baseline: linear sample done ({wall:.1f} s)")
    return {"solves_per_s": len(out) / wall, "threads": cores, "gamma": 0.01,
            "cg_iter_cap": 20000, "networks": out}


if __name__ == "__main__":
    main()
