"""GPU tests of the device strain ramp and the serving loop, through the C ABI.

A ramp walked by the kernel (nr_batch_ramp) must give bit-identical reports
and positions to the same increments applied one relax call at a time; a
served job, paused and resumed every few dozen steps, must be bit-identical
to the same job relaxed in one launch, alone or in any mix; and the first
increments are checked against the FP64 oracle (1e-12 relative energy).
"""
import sys

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

pytestmark = pytest.mark.gpu

PRM = dict(dt_max=0.2, fcut=1e-10)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _yrange(net):
    i = net.info()
    return i.height if i.height > 0 else i.maxy - i.miny


def _incs(net, shear, comp=()):
    """Protocol-style increments: compression steps, then shear steps
    (strain mode: displacement = strain x y range)."""
    yr = _yrange(net)
    return np.array([(0.0, c * yr) for c in comp] + [(g * yr, 0.0) for g in shear])


def _key(r):
    return (r["steps"], r["status"], r["es"], r["eb"], r["fmax"])


def test_ramp_equals_sequential_relaxes_and_oracle():
    _say("case ramp vs sequential")
    specs = [(20, 0.7, 4, True), (24, 0.68, 9, False), (16, 0.72, 2, True)]
    prm = N.relax_params(**PRM)
    nets = [N.Network.lattice(L, p, s, le).triples() for L, p, s, le in specs]
    shear = [0.01, 0.01, 0.015, 0.01]
    incs = [_incs(n, shear, comp=() if specs[k][3] else (0.005,)) for k, n in enumerate(nets)]
    # one network per batch (increments differ between LE and clamped)
    for k, n in enumerate(nets):
        rb = N.Batch([n], 1.0, 5e-3)
        ramp = rb.ramp(incs[k], prm)[0]
        sb = N.Batch([n], 1.0, 5e-3)
        for i, (dx, dy) in enumerate(incs[k]):
            sb.strain(0, dx, dy)
            r = sb.relax(prm)[0]
            assert _key(r) == _key(ramp[i]), (k, i)
        np.testing.assert_array_equal(rb.positions(0), sb.positions(0))
    # the first network against the oracle, increment by increment
    L, p, s, le = specs[0]
    o = O.Net.lattice(L, p, s, le).triples()
    pos = o.pos()
    rb = N.Batch([nets[0]], 1.0, 5e-3)
    ramp = rb.ramp(incs[0], prm)[0]
    for i, g in enumerate(shear):
        o.shear(pos, g)
        ro = o.fire(1.0, 5e-3, pos, fp=(0.1, 0.99, 0.2, 1.1, 0.5), fcut=1e-10)
        et, eo = ramp[i]["es"] + ramp[i]["eb"], ro["es"] + ro["eb"]
        assert abs(et - eo) <= 1e-12 * eo, (i, et, eo)


def _served(jobs, precision=N.PREC_FP64, slice_steps=40, hbm=0):
    """Submits (net, kb, incs) jobs to a fresh server and drains it."""
    srv = N.Server(precision=precision, slice_steps=slice_steps, hbm_bytes=hbm)
    prm = N.relax_params(**PRM)
    ids = [srv.submit(n, 1.0, kb, prm, incs, keep_positions=True) for n, kb, incs in jobs]
    srv.drain()
    res = {r["job"]: r for r in srv.results()}
    info = srv.stats()
    assert info["nodes_free"] == info["node_cap"] and info["active"] == 0
    return [res[i] for i in ids], info


@pytest.mark.parametrize("precision", [N.PREC_FP64, N.PREC_MIXED])
def test_served_job_equals_one_launch_ramp(precision):
    """Paused every 40 steps and resumed: bit-identical to one launch."""
    _say(f"case served vs ramp precision={precision}")
    mk = lambda L, s: N.Network.lattice(L, 0.66, s, True).triples().fiber_kappa(1e-4, 1e-2, s)
    jobs = []
    for L, s, ninc in ((16, 1, 3), (40, 2, 2), (28, 3, 1), (64, 4, 2)):
        n = mk(L, s)
        jobs.append((n, 1.0, _incs(n, [0.02] * ninc)))
    res, info = _served(jobs, precision=precision)
    assert info["slices"] > 3  # the jobs really were paused and resumed
    prm = N.relax_params(**PRM)
    for (n, kb, incs), r in zip(jobs, res):
        assert r["error"] == N.OK and r["n_inc"] == len(incs)
        # same shape as the server's (one tile per team member): same arithmetic
        b = N.Batch([n], 1.0, kb, precision=precision, tiles_per_member=1)
        ramp = b.ramp(incs, prm)[0]
        assert [_key(x) for x in r["reports"]] == [_key(x) for x in ramp]
        np.testing.assert_array_equal(r["pos"], b.positions(0))
        # the reductions the server attaches equal the batch's of the same state
        yr = n.info().height
        gx = gy = 0.0
        for dx, dy in incs:
            gx += dx
            gy += dy
        a = b.analyze(0, gx / yr, gy / yr, 1.0)
        assert (r["nonaffinity"], r["axial_energy"], r["mean_bond_strain"],
                r["max_bond_strain"]) == (a["nonaffinity"], a["axial_energy"],
                                          a["mean_strain"], a["max_strain"])
        assert r["nonaffinity"] > 0.0


def test_served_alone_equals_in_mix():
    """Each job's result is bit-identical alone and in a mix of uneven jobs
    (mixed precision, small slices, memory budget forcing queueing)."""
    _say("case served alone vs mix")
    mk = lambda L, s: N.Network.lattice(L, 0.64, 50 + s, True).triples().fiber_kappa(1e-5, 1e-2, s)
    specs = [(18, 1, 2), (72, 2, 1), (24, 3, 3), (48, 4, 2), (16, 5, 1), (96, 6, 1)]
    jobs = [(mk(L, s), 1.0, None) for L, s, _ in specs]
    jobs = [(n, kb, _incs(n, [0.03] * specs[k][2])) for k, (n, kb, _) in enumerate(jobs)]
    # 12000 nodes of arena (233 B each when mixed): the two largest jobs
    # (9216 and 5184 nodes) cannot be resident together
    mix, info = _served(jobs, precision=N.PREC_MIXED, slice_steps=60, hbm=12000 * 233)
    for k, job in enumerate(jobs):
        alone, _ = _served([job], precision=N.PREC_MIXED, slice_steps=60)
        assert [_key(x) for x in alone[0]["reports"]] == [_key(x) for x in mix[k]["reports"]], k
        np.testing.assert_array_equal(alone[0]["pos"], mix[k]["pos"])
    # continuous batching: some job was admitted while another was resident
    overlap = any(a["t_admit"] > b["t_admit"] and a["t_admit"] < b["t_done"]
                  for a in mix for b in mix if a is not b)
    assert overlap


def test_server_rejects_job_larger_than_arena():
    _say("case server too large")
    small = N.Network.lattice(16, 0.7, 1, True).triples()
    big = N.Network.lattice(128, 0.7, 2, True).triples()
    srv = N.Server(hbm_bytes=4000 * 233)
    prm = N.relax_params(**PRM)
    jb = srv.submit(big, 1.0, 1e-3, prm, _incs(big, [0.01]))
    js = srv.submit(small, 1.0, 1e-3, prm, _incs(small, [0.01]))
    srv.drain()
    res = {r["job"]: r for r in srv.results()}
    assert res[jb]["status"] == N.ABORTED and res[jb]["error"] == N.E_ALLOC
    assert res[js]["status"] == N.CONVERGED and res[js]["n_inc"] == 1


def test_multi_member_teams_alone_equal_in_mix():
    """Networks of 9 to 64 tiles (teams of 2 to 16 workgroups): bit-identical
    alone and in one batch.  Members read the network's records at the
    start of a relax while its lead writes them at the end, so any write
    that came earlier would show here as a different result."""
    _say("case multi-member teams alone vs mix")
    sizes = (48, 64, 80, 96, 112, 128)
    mk = lambda k: N.Network.lattice(sizes[k], 0.66, 70 + k, True).triples()
    prm = N.relax_params(dt_max=0.3, fcut=1e-9)
    mix = N.Batch([mk(k) for k in range(len(sizes))], 1.0, 2e-3, tiles_per_member=4)
    mix.shear(-1, 0.04)
    rm = mix.relax(prm)
    for k in range(len(sizes)):
        alone = N.Batch([mk(k)], 1.0, 2e-3, tiles_per_member=4)
        alone.shear(0, 0.04)
        ra = alone.relax(prm)[0]
        assert _key(ra) == _key(rm[k]), k
        np.testing.assert_array_equal(alone.positions(0), mix.positions(k))


def test_graph_step_replays_equal_eager_steps():
    """A captured strain step (device affine step + relax) replayed three
    times, with eager work and synchronisation between the replays, equals
    the same three steps taken eagerly, bit for bit, every time."""
    _say("case graph replay")
    sizes = (16, 40, 72)
    mk = lambda k: N.Network.lattice(sizes[k], 0.7, 90 + k, k != 1).triples()
    nets_g = [mk(k) for k in range(3)]
    nets_e = [mk(k) for k in range(3)]
    prm = N.relax_params(dt_max=0.2, fcut=1e-10)
    yr = np.array([_yrange(n) for n in nets_g])
    xd = 0.01 * yr
    yd = np.array([0.0, -0.002 * yr[1], 0.0])  # compression on the clamped network
    g = N.Batch(nets_g, 1.0, 4e-3)
    e = N.Batch(nets_e, 1.0, 4e-3)
    g.graph_capture(xd, yd, prm)
    for step in range(3):
        rg = g.graph_replay()
        for k in range(3):
            e.strain(k, xd[k], yd[k])
        re = e.relax(prm)
        assert [_key(x) for x in rg] == [_key(x) for x in re], step
        for k in range(3):  # eager work between replays
            np.testing.assert_array_equal(g.positions(k), e.positions(k))


def test_autotune_returns_a_working_shape():
    """The tuner's choice is a valid shape whose relax reaches the same
    minima (team sizes change only the reduction order)."""
    _say("case autotune")
    sizes = (16, 48, 96, 64, 24)
    nets = [N.Network.lattice(L, 0.66, 200 + L, True).triples() for L in sizes]
    shape, pred = N.autotune(nets, 1.0, 2e-3, probe_steps=50)
    assert shape["block"] in (256, 512) and shape["tiles_per_member"] >= 1 and pred > 0
    prm = N.relax_params(dt_max=0.2, fcut=1e-10)
    tuned = N.Batch(nets, 1.0, 2e-3, **shape)
    base = N.Batch(nets, 1.0, 2e-3)
    tuned.shear(-1, 0.01)
    base.shear(-1, 0.01)
    rt, rb = tuned.relax(prm), base.relax(prm)
    for a, b in zip(rt, rb):
        assert a["status"] == N.CONVERGED
        ea, eb = a["es"] + a["eb"], b["es"] + b["eb"]
        assert abs(ea - eb) <= 1e-12 * eb
