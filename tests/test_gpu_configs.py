"""GPU: the relax path at the full sizes of the stated workloads, each result
checked by the FP64 oracle evaluating forces and energies on the device's
returned positions (the oracle computes those from the reference's kernels,
large_strain_relax.cu:386-523).

  L=512 ramp      L=512, p 0.6, kappa 1e-4, Lees-Edwards box: three shear
                  increments of 0.005 from the unstrained network, each relaxed
                  from the last (do_relaxation_run's loop, :966-999), by FIRE
                  and by nonlinear CG.
  L=2048 mixed    L=2048, p 0.62, per-fiber kappa log-U[1e-5, 1e-2]: one shear
                  increment of 0.005, mixed precision against FP64.
  512 batch       the 512-network batch (L even U[16,192], p U[0.55,0.75],
                  per-fiber kappa log-U[1e-5,1e-2], gamma U[0,0.4]): every
                  network, mixed precision and FP64; a sample relaxed alone.
  served stream   jobs with L log-U[16,1024] (one L=1024 job), p U[0.55,0.75],
                  1-4 shear increments of U[0.01,0.05] through the serving
                  loop; a sample in FP64 and a sample alone.

Pass criteria, per relaxed state: max node |F| of the oracle's own force
evaluation below the cut, and the reported energy equal to the oracle's
within 1e-12 relative; mixed precision within 1e-10 relative of FP64; results
in a batch or in the server's mix bit-identical to the same network relaxed
alone (team sizes fixed by tiles_per_member).
"""
import sys
import time

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

FIRE = dict(alpha0=0.1, f_alpha=0.99, dt_max=0.5, f_inc=1.1, f_dec=0.5, nmin=5, mass=1.0)
FCUT = 1e-9


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _shift_after(incs):
    """Lees-Edwards image shift after the shear steps incs, added in the
    order the device adds them."""
    s = 0.0
    for d in incs:
        s += d
    return s


def _oracle_check(o, pos, shift, ks, kb, rep, fcut=FCUT, what=""):
    """The oracle's FP64 forces and energies at the device's positions."""
    o.s.le_shift = shift
    f = o.forces(ks, kb, pos).reshape(-1, 2)
    fmax = float(np.sqrt((f * f).sum(axis=1)).max()) if len(f) else 0.0
    es, eb = o.energies(ks, kb, pos)
    e_dev, e_or = rep["es"] + rep["eb"], es + eb
    assert fmax < fcut, f"{what}: oracle max|F| {fmax:.3e} at the device's state"
    assert abs(e_dev - e_or) <= 1e-12 * abs(e_or), \
        f"{what}: energy {e_dev!r} vs oracle {e_or!r}"
    assert abs(rep["es"] - es) <= 1e-12 * abs(e_or) and abs(rep["eb"] - eb) <= 1e-12 * abs(e_or)
    return fmax, e_or


# ---- L=512 shear ramp, FIRE and CG --------------------------------------

@pytest.fixture(scope="module")
def ramp512():
    L, p, seed, kb = 512, 0.6, 2, 1e-4
    n = N.Network.lattice(L, p, seed, True).triples()
    d = 0.005 * n.info().height
    incs = np.array([(d, 0.0)] * 3)
    out = {"net": n, "incs": incs, "kb": kb}
    for name, method in (("fire", N.METHOD_FIRE), ("cg", N.METHOD_CG)):
        _say(f"case L=512 ramp {name}: 3 x 0.005")
        b = N.Batch([n], 1.0, kb)
        prm = N.relax_params(fcut=FCUT, criterion=N.CRIT_MAX, method=method,
                             max_steps=2000000, **FIRE)
        t = time.perf_counter()
        reps = b.ramp(incs, prm)[0]
        _say(f"  {name}: steps {[r['steps'] for r in reps]} in {time.perf_counter() - t:.1f} s")
        out[name] = (reps, b.positions(0))
        del b
    return out


def test_ramp512_fire_and_cg_against_oracle(ramp512):
    o = O.Net.lattice(512, 0.6, 2, True).triples()
    shift = _shift_after(ramp512["incs"][:, 0])
    for name in ("fire", "cg"):
        reps, pos = ramp512[name]
        assert all(r["status"] == N.CONVERGED for r in reps), [r["status"] for r in reps]
        fmax, e = _oracle_check(o, pos, shift, 1.0, ramp512["kb"], reps[-1], what=name)
        _say(f"  {name}: oracle max|F| {fmax:.2e}, E {e:.15e}")


def test_ramp512_fire_and_cg_reach_one_minimum(ramp512):
    """Both minimisers start the first increment from the same affine state
    and end in its basin.  The network is floppy (p = 0.6 is below the
    rigidity threshold; only kappa = 1e-4 bending holds it), so a state with
    max |F| < 1e-9 can still sit |F|^2 / k_soft above the minimum along a
    soft mode: measured 3.5e-8 relative between the two at the first
    increment.  Energies must agree to 1e-7 there and 1e-6 at the third
    increment (different minima of a multistable network would differ by
    1e-5 and more, as found after 20 increments)."""
    rf, rc = ramp512["fire"][0], ramp512["cg"][0]
    for i, tol in ((0, 1e-7), (2, 1e-6)):
        ef, ec = rf[i]["es"] + rf[i]["eb"], rc[i]["es"] + rc[i]["eb"]
        _say(f"  increment {i + 1}: FIRE {ef:.15e} CG {ec:.15e} rel {abs(ef - ec) / ef:.2e}")
        assert abs(ef - ec) <= tol * ef


# ---- L=2048, mixed precision against FP64 --------------------------------

@pytest.fixture(scope="module")
def large2048():
    n = N.Network.lattice(2048, 0.62, 3, True).triples().fiber_kappa(1e-5, 1e-2, 3)
    d = 0.005 * n.info().height
    return {"net": n, "incs": np.array([(d, 0.0)]), "res": {}}


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_large2048_against_oracle(large2048, prec):
    _say(f"case L=2048 one increment of 0.005, {prec}")
    n = large2048["net"]
    b = N.Batch([n], 1.0, 1.0, precision=N.PREC_MIXED if prec == "mixed" else N.PREC_FP64)
    prm = N.relax_params(fcut=FCUT, criterion=N.CRIT_MAX, max_steps=3000000, **FIRE)
    t = time.perf_counter()
    rep = b.ramp(large2048["incs"], prm)[0][0]
    _say(f"  {prec}: {rep['steps']} steps ({rep['fp64_steps']} FP64) in "
         f"{time.perf_counter() - t:.1f} s, sxy {rep['sxy']:.15e}")
    assert rep["status"] == N.CONVERGED
    pos = b.positions(0)
    del b
    o = O.Net.lattice(2048, 0.62, 3, True).triples().fiber_kappa(1e-5, 1e-2, 3)
    shift = _shift_after(large2048["incs"][:, 0])
    fmax, e = _oracle_check(o, pos, shift, 1.0, 1.0, rep, what=prec)
    # shear stress: the oracle's virial at the same state over the box area
    o.s.le_shift = shift
    sxy = o.shear_virial(1.0, 1.0, pos) / (n.info().width * n.info().height)
    assert abs(rep["sxy"] - sxy) <= 1e-10 * abs(sxy)
    large2048["res"][prec] = rep


def test_large2048_mixed_within_1e10_of_fp64(large2048):
    res = large2048["res"]
    if set(res) != {"mixed", "fp64"}:
        pytest.skip("needs both L=2048 relaxes")
    em = res["mixed"]["es"] + res["mixed"]["eb"]
    ef = res["fp64"]["es"] + res["fp64"]["eb"]
    _say(f"  L=2048 mixed vs FP64: energy rel {abs(em - ef) / ef:.2e}, stress rel "
         f"{abs(res['mixed']['sxy'] - res['fp64']['sxy']) / abs(res['fp64']['sxy']):.2e}")
    assert abs(em - ef) <= 1e-10 * ef
    assert abs(res["mixed"]["sxy"] - res["fp64"]["sxy"]) <= 1e-7 * abs(res["fp64"]["sxy"])


# ---- the 512-network batch ------------------------------------------------

def _batch_specs(n_nets=512, seed=2024):
    """The batch workload: L even U[16,192], p U[0.55,0.75], gamma U[0,0.4],
    per-fiber kappa log-U[1e-5,1e-2] (the benchmark's seeds)."""
    rng = np.random.RandomState(seed)
    specs = []
    for i in range(n_nets):
        L = 2 * rng.randint(8, 97)
        p = rng.uniform(0.55, 0.75)
        gamma = rng.uniform(0.0, 0.4)
        specs.append(dict(L=int(L), p=float(p), gamma=float(gamma), seed=1000 + i, kseed=5000 + i))
    return specs


def _batch_net(s):
    return N.Network.lattice(s["L"], s["p"], s["seed"], True).triples().fiber_kappa(
        1e-5, 1e-2, s["kseed"])


@pytest.fixture(scope="module")
def batch512():
    specs = _batch_specs()
    t = time.perf_counter()
    nets = [_batch_net(s) for s in specs]
    _say(f"case 512-network batch: built in {time.perf_counter() - t:.1f} s, "
         f"{sum(s['L'] ** 2 for s in specs)} nodes")
    prm = N.relax_params(fcut=FCUT, criterion=N.CRIT_MAX, max_steps=3000000, **FIRE)
    out = {"specs": specs, "nets": nets, "prm": prm}
    for name, prec in (("mixed", N.PREC_MIXED), ("fp64", N.PREC_FP64)):
        b = N.Batch(nets, 1.0, 1.0, precision=prec, tiles_per_member=8)
        for k, s in enumerate(specs):
            b.shear(k, s["gamma"])
        t = time.perf_counter()
        reps = b.relax(prm)
        _say(f"  {name}: relaxed in {time.perf_counter() - t:.1f} s "
             f"(kernel {b.last_ms():.0f} ms)")
        out[name] = (reps, [b.positions(k) for k in range(len(specs))])
        del b
    return out


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_batch512_every_network_against_oracle(batch512, prec):
    reps, pos = batch512[prec]
    assert all(r["status"] == N.CONVERGED for r in reps)
    worst = 0.0
    for k, s in enumerate(batch512["specs"]):
        o = O.Net.lattice(s["L"], s["p"], s["seed"], True).triples().fiber_kappa(
            1e-5, 1e-2, s["kseed"])
        info = batch512["nets"][k].info()
        shift = _shift_after([s["gamma"] * info.height])
        fmax, _ = _oracle_check(o, pos[k], shift, 1.0, 1.0, reps[k], what=f"{prec} net {k}")
        worst = max(worst, fmax)
    _say(f"  {prec}: 512 networks, oracle max|F| at most {worst:.2e}")


def test_batch512_mixed_within_1e10_of_fp64(batch512):
    em = np.array([r["es"] + r["eb"] for r in batch512["mixed"][0]])
    ef = np.array([r["es"] + r["eb"] for r in batch512["fp64"][0]])
    rel = np.abs(em - ef) / np.abs(ef)
    _say(f"  mixed vs FP64: median {np.median(rel):.2e} max {rel.max():.2e}")
    assert (rel <= 1e-10).all(), np.flatnonzero(rel > 1e-10)


def test_batch512_alone_equals_in_batch(batch512):
    """tiles_per_member fixed: a network's team size, and so its arithmetic,
    does not depend on the rest of the batch."""
    specs = batch512["specs"]
    order = np.argsort([s["L"] for s in specs])
    sample = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    reps, pos = batch512["mixed"]
    for k in sample:
        s = specs[k]
        b = N.Batch([_batch_net(s)], 1.0, 1.0, precision=N.PREC_MIXED, tiles_per_member=8)
        b.shear(0, s["gamma"])
        r = b.relax(batch512["prm"])[0]
        _say(f"  alone: net {k} L={s['L']} {r['steps']} steps")
        assert (r["steps"], r["fp64_steps"], r["es"], r["eb"]) == \
            (reps[k]["steps"], reps[k]["fp64_steps"], reps[k]["es"], reps[k]["eb"])
        np.testing.assert_array_equal(b.positions(0), pos[k])


# ---- served stream, L log-U[16,1024] --------------------------------------

def _served_jobs(n_jobs=10, seed=77):
    """Jobs with L log-uniform in [16, 1024] (L even; the first L = 1024),
    p U[0.55,0.75], per-fiber kappa log-U[1e-5,1e-2], 1-4 shear increments
    of U[0.01,0.05] (2 for the L = 1024 job)."""
    rng = np.random.RandomState(seed)
    jobs = []
    for i in range(n_jobs):
        L = int(round(np.exp(rng.uniform(np.log(16), np.log(1024))) / 2)) * 2
        p = float(rng.uniform(0.55, 0.75))
        ninc = int(rng.randint(1, 5))
        if i == 0:
            L, ninc = 1024, 2
        gam = rng.uniform(0.01, 0.05, size=ninc)
        jobs.append(dict(L=max(L, 16), p=p, seed=9000 + i, kseed=7000 + i, gammas=gam))
    return jobs


def _job_net(j):
    return N.Network.lattice(j["L"], j["p"], j["seed"], True).triples().fiber_kappa(
        1e-5, 1e-2, j["kseed"])


def _serve(jobs, nets, precision, prm):
    srv = N.Server(precision=precision, slice_steps=2000)
    ids = {}
    for k, (j, n) in enumerate(zip(jobs, nets)):
        h = n.info().height
        incs = np.stack([j["gammas"] * h, np.zeros(len(j["gammas"]))], axis=1)
        ids[srv.submit(n, 1.0, 1.0, prm, incs, keep_positions=True)] = k
    res = {}
    last = time.perf_counter()
    while len(res) < len(jobs):
        srv.step()
        for r in srv.results():
            res[ids[r["job"]]] = r
        if time.perf_counter() - last > 15:
            last = time.perf_counter()
            _say(f"  served: {len(res)}/{len(jobs)} done, {srv.stats()['slices']} launches")
    return res


@pytest.fixture(scope="module")
def served():
    jobs = _served_jobs()
    nets = [_job_net(j) for j in jobs]
    _say(f"case served stream: L {[j['L'] for j in jobs]}, increments "
         f"{[len(j['gammas']) for j in jobs]}")
    prm = N.relax_params(fcut=FCUT, criterion=N.CRIT_MAX, max_steps=3000000, **FIRE)
    t = time.perf_counter()
    res = _serve(jobs, nets, N.PREC_MIXED, prm)
    _say(f"  mixed stream done in {time.perf_counter() - t:.1f} s")
    return {"jobs": jobs, "nets": nets, "prm": prm, "mixed": res}


def test_served_every_job_against_oracle(served):
    for k, j in enumerate(served["jobs"]):
        r = served["mixed"][k]
        assert r["status"] == N.CONVERGED and r["n_inc"] == len(j["gammas"])
        o = O.Net.lattice(j["L"], j["p"], j["seed"], True).triples().fiber_kappa(
            1e-5, 1e-2, j["kseed"])
        h = served["nets"][k].info().height
        shift = _shift_after(list(j["gammas"] * h))
        fmax, _ = _oracle_check(o, r["pos"], shift, 1.0, 1.0, r["reports"][-1],
                                what=f"job {k} L={j['L']}")
        _say(f"  job {k} L={j['L']}: {r['steps']} steps, oracle max|F| {fmax:.2e}")


def test_served_mixed_within_1e10_of_fp64(served):
    """The L=1024 job and two more, in an FP64 server: every increment's
    energy within 1e-10 relative of the mixed stream's."""
    pick = [0, 1, 2]
    res = _serve([served["jobs"][k] for k in pick], [served["nets"][k] for k in pick],
                 N.PREC_FP64, served["prm"])
    for i, k in enumerate(pick):
        em = np.array([x["es"] + x["eb"] for x in served["mixed"][k]["reports"]])
        ef = np.array([x["es"] + x["eb"] for x in res[i]["reports"]])
        rel = np.abs(em - ef) / np.abs(ef)
        _say(f"  job {k}: mixed vs FP64 per increment {rel}")
        assert (rel <= 1e-10).all()


def test_served_alone_equals_in_mix(served):
    """Two jobs re-run alone in a fresh server: reports and final positions
    bit-identical to the mixed stream's."""
    for k in (1, 2):
        res = _serve([served["jobs"][k]], [_job_net(served["jobs"][k])], N.PREC_MIXED,
                     served["prm"])[0]
        mix = served["mixed"][k]
        assert [(x["steps"], x["es"], x["eb"]) for x in res["reports"]] == \
               [(x["steps"], x["es"], x["eb"]) for x in mix["reports"]]
        np.testing.assert_array_equal(res["pos"], mix["pos"])
