"""The network + FEM-gel composite on the device (nr_composite_*, the
nrelax_mesh CLI) against the oracle's restatement of
src/network_mesh_relax.cu (oracle/mesh.c, pinned in tests/test_mesh.py).

Tolerances: the set-up runs the same host arithmetic as the oracle (relative
1e-15); forces agree to 1e-12 of the largest force (the device sums each
point's terms in another order); relaxed energies to 1e-10 of the total at
|F|_rms < 1e-11 (two FIRE trajectories that differ in the last bits end within
O(fcut^2 / stiffness) of one minimum).
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

pytestmark = pytest.mark.gpu

FP = (0.1, 0.99, 0.1, 1.1, 0.5)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(tmp_path, L, p, seed, ks, kb, bmod, smod, dil, name="c"):
    path = os.path.join(str(tmp_path), f"{name}.nmc")
    mesh = N.Mesh.lattice(L, p, seed)
    mesh.write_nmc(path)
    c = N.Composite(mesh, ks, kb, bmod, smod, dil)
    o = O.Mesh.read(path)
    pos, pd = o.setup(ks, kb, bmod, dil)
    return c, o, pos, pd


@pytest.mark.parametrize("dil", [0.0, 0.1])
def test_setup_matches_oracle(tmp_path, dil):
    _say(f"case setup dil={dil}")
    c, o, pos, pd = _pair(tmp_path, 16, 0.7, 3, 1.0, 1e-2, 0.8, 0.3, dil)
    i = c.info()
    assert i.pref_d == pytest.approx(pd, rel=1e-15, abs=0.0)
    assert i.n_triples == o.t and i.n_bonds == o.e and i.n_facets == o.nf
    assert i.width == pytest.approx(16 * math.sqrt(1 + dil), rel=1e-15)
    np.testing.assert_allclose(c.positions(), pos, rtol=0, atol=1e-15)


@pytest.mark.parametrize("ks,kb,bmod,smod", [(1.0, 2e-2, 0.9, 0.4), (1.0, 2e-2, 0.0, 0.4),
                                             (1.0, 0.0, 0.9, 0.4), (0.0, 0.0, 0.9, 0.4)])
def test_forces_and_energies_match_oracle(tmp_path, ks, kb, bmod, smod):
    _say(f"case eval ks={ks} kb={kb} bmod={bmod}")
    c, o, pos, pd = _pair(tmp_path, 12, 0.75, 4, ks, kb, bmod, smod, 0.08)
    rng = np.random.default_rng(5)
    for state in (pos, pos + 0.02 * rng.standard_normal(pos.size)):
        c.set_positions(state)
        f, e = c.eval()
        fo = o.forces(ks, kb, smod, bmod, pd, state)
        eo = o.energies(ks, kb, smod, bmod, pd, state)
        scale = max(1.0, np.max(np.abs(fo)))
        assert np.max(np.abs(f - fo)) <= 1e-12 * scale
        np.testing.assert_allclose(e, eo, rtol=1e-12, atol=1e-14)
        if bmod > 0:
            for a, b in o.pairs():
                assert f[2 * a] == f[2 * b] and f[2 * a + 1] == f[2 * b + 1]


@pytest.mark.parametrize("L,dil,gamma", [(16, 0.05, 0.05), (32, 0.1, 0.02)])
def test_relax_matches_oracle(tmp_path, L, dil, gamma):
    _say(f"case relax L={L} dil={dil} gamma={gamma}")
    ks, kb, bmod, smod = 1.0, 1e-3, 0.5, 0.25
    c, o, pos, pd = _pair(tmp_path, L, 0.6, 11, ks, kb, bmod, smod, dil)
    yr = c.info().range
    c.affine(yr, gamma * yr, 0.0)
    o.affine(pos, gamma * yr, 0.0, yr)
    prm = N.relax_params(fcut=1e-11, criterion=N.CRIT_RMS)
    r = c.relax(prm)
    ro = o.fire(ks, kb, smod, bmod, pd, pos, fcut=1e-11)
    assert r["status"] == N.CONVERGED and ro["rc"] == 0
    assert r["rms"] < 1e-11
    tot = ro["es"] + ro["eb"] + ro["eg"]
    for k in ("es", "eb", "eg"):
        assert abs(r[k] - ro[k]) <= 1e-10 * tot, k
    assert abs(r["steps"] - ro["steps"]) <= 0.05 * ro["steps"] + 5
    np.testing.assert_allclose(c.positions(), pos, rtol=0, atol=1e-7)


def test_sheared_gel_equilibrium_closed_form(tmp_path):
    """Gel only, no dilation: the affine shear is the minimum (one FIRE step)
    and E_gel = smod A gamma^2 / 2."""
    _say("case gel-only shear")
    c, o, pos, pd = _pair(tmp_path, 20, 0.6, 6, 0.0, 0.0, 1.0, 0.7, 0.0)
    yr = c.info().range
    g = 0.05
    c.affine(yr, g * yr, 0.0)
    r = c.relax(N.relax_params(fcut=1e-10, criterion=N.CRIT_RMS))
    assert r["status"] == N.CONVERGED and r["steps"] == 1
    assert r["eg"] == pytest.approx(0.7 * o.areas().sum() * g * g / 2, rel=1e-10)
    assert r["es"] == 0.0 and r["eb"] == 0.0


def test_zero_step_limit_leaves_positions(tmp_path):
    _say("case max_steps 0")
    c, o, pos, pd = _pair(tmp_path, 10, 0.7, 2, 1.0, 1e-2, 0.5, 0.2, 0.05)
    yr = c.info().range
    c.affine(yr, 0.02 * yr, 0.0)
    before = c.positions()
    r = c.relax(N.relax_params(fcut=1e-10, max_steps=0, criterion=N.CRIT_RMS))
    assert r["status"] == N.NOT_CONVERGED and r["steps"] == 0
    np.testing.assert_array_equal(c.positions(), before)


def test_ramp_keeps_state_between_increments(tmp_path):
    """Compression then two shear increments, positions resident on the
    device, each relax checked against the oracle doing the same steps."""
    _say("case ramp")
    ks, kb, bmod, smod = 1.0, 1e-3, 0.5, 0.25
    c, o, pos, pd = _pair(tmp_path, 16, 0.65, 9, ks, kb, bmod, smod, 0.05)
    yr = c.info().range
    prm = N.relax_params(fcut=1e-11, criterion=N.CRIT_RMS)
    steps = [(0.0, -0.01 * yr, yr), (0.02 * yr, 0.0, yr - 0.01 * yr),
             (0.02 * yr, 0.0, yr - 0.01 * yr)]
    for xd, ydisp, rng_ in steps:
        c.affine(rng_, xd, ydisp)
        o.affine(pos, xd, ydisp, rng_)
        r = c.relax(prm)
        ro = o.fire(ks, kb, smod, bmod, pd, pos, fcut=1e-11)
        tot = ro["es"] + ro["eb"] + ro["eg"]
        assert r["status"] == N.CONVERGED
        assert abs((r["es"] + r["eb"] + r["eg"]) - tot) <= 1e-10 * tot


def test_mesh_cli_matches_oracle(tmp_path):
    """nrelax_mesh -b runs the composite program's protocol: the same files
    as the oracle's run, energies within 1e-9, positions within 1e-7, the
    preferred dilation on stdout."""
    _say("case nrelax_mesh CLI")
    d = str(tmp_path)
    os.makedirs(f"{d}/gpu")
    os.makedirs(f"{d}/cpu")
    for sub in ("gpu", "cpu"):
        N.Mesh.lattice(10, 0.7, 8).write_nmc(f"{d}/{sub}/cmp.nmc")
        with open(f"{d}/{sub}/b.txt", "w") as f:
            f.write(f"{d}/{sub}/cmp t\n1.0 0.01 0.8 0.4 0.05\n0.1 0.99 0.1 1.1 0.5\n"
                    "5 1.0 1e-11 500000\n0 0.02 0.01\n0.01 0.03 0.01\n")
    out = subprocess.run([N.MESH_CLI_PATH, "-b", f"{d}/gpu/b.txt"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rc, _, _, b = O.parse_mesh_batch(f"{d}/cpu/b.txt")
    rc, pd = O.run_mesh_protocol(f"{d}/cpu/cmp.nmc", f"{d}/cpu/cmp_t_log.txt",
                                 f"{d}/cpu/cmp_t_edata.txt", f"{d}/cpu/cmp_t", b, True)
    assert rc == 0
    assert out.stdout.splitlines()[0] == f"Preferred dilation: {pd:g}"
    eg = open(f"{d}/gpu/cmp_t_edata.txt").read().splitlines()
    ec = open(f"{d}/cpu/cmp_t_edata.txt").read().splitlines()
    assert len(eg) == len(ec) == 6
    for a, b_ in zip(eg, ec):
        fa = [float(x) for x in a.split("\t")]
        fb = [float(x) for x in b_.split("\t")]
        assert a.split("\t")[:2] == b_.split("\t")[:2]
        np.testing.assert_allclose(fa[2:], fb[2:], rtol=1e-9, atol=1e-15)
    lg = open(f"{d}/gpu/cmp_t_log.txt").read().splitlines()
    lc = open(f"{d}/cpu/cmp_t_log.txt").read().splitlines()
    assert len(lg) == len(lc) == 7
    assert all(x.startswith("Convergence reached: Steps: ") for x in lg[:6])
    assert lg[6].startswith("Elapsed time: ") and lg[6].endswith(" seconds")
    for k in (1, 2, 3):
        pg = np.loadtxt(f"{d}/gpu/cmp_t_{k}.pos")
        pc = np.loadtxt(f"{d}/cpu/cmp_t_{k}.pos")
        np.testing.assert_allclose(pg, pc, rtol=0, atol=1e-7)


def test_direct_launches_equal_graph_replays(tmp_path, monkeypatch):
    """NRELAX_GRAPHS=0 at creation launches the same kernels directly: the
    same steps, energies and positions, bit for bit."""
    _say("case direct launches vs graph replays")
    mesh = N.Mesh.lattice(24, 0.65, 5)
    out = []
    for env in ("1", "0"):
        monkeypatch.setenv("NRELAX_GRAPHS", env)
        c = N.Composite(mesh, 1.0, 1e-3, 0.5, 0.25, 0.05)
        yr = c.info().range
        c.affine(yr, 0.03 * yr, 0.0)
        r = c.relax(N.relax_params(fcut=1e-11, criterion=N.CRIT_RMS))
        out.append((r["steps"], r["es"], r["eb"], r["eg"], c.positions()))
    assert out[0][:4] == out[1][:4]
    np.testing.assert_array_equal(out[0][4], out[1][4])


@pytest.mark.parametrize("L,dil", [(24, 0.05), (96, 0.1)])
def test_tiled_step_equals_element_and_gather_kernels(tmp_path, monkeypatch, L, dil):
    """The fused tile kernel (the default) and the unfused element + gather
    kernels (NRELAX_MESH_TILES=0 at creation) give the same forces bit for
    bit on a perturbed state (same element arithmetic, same entry order per
    point), and relaxes that agree to the relax tolerance (the partials are
    summed in another order)."""
    _say(f"case tiled vs unfused L={L}")
    mesh = N.Mesh.lattice(L, 0.62, 8)
    cs = {}
    for env in ("1", "0"):
        monkeypatch.setenv("NRELAX_MESH_TILES", env)
        cs[env] = N.Composite(mesh, 1.0, 1e-3, 0.5, 0.25, dil)
    start = cs["1"].positions()
    np.testing.assert_array_equal(start, cs["0"].positions())
    state = start + 0.02 * np.random.default_rng(11).standard_normal(start.size)
    f = {}
    for env, c in cs.items():
        c.set_positions(state)
        f[env] = c.eval()
    np.testing.assert_array_equal(f["1"][0], f["0"][0])
    np.testing.assert_allclose(f["1"][1], f["0"][1], rtol=1e-13, atol=0.0)
    res = {}
    for env, c in cs.items():
        c.set_positions(start)
        yr = c.info().range
        c.affine(yr, 0.03 * yr, 0.0)
        res[env] = c.relax(N.relax_params(fcut=1e-11, criterion=N.CRIT_RMS))
    e1 = res["1"]["es"] + res["1"]["eb"] + res["1"]["eg"]
    e0 = res["0"]["es"] + res["0"]["eb"] + res["0"]["eg"]
    assert res["1"]["status"] == N.CONVERGED and res["0"]["status"] == N.CONVERGED
    assert abs(e1 - e0) <= 1e-10 * abs(e0)
