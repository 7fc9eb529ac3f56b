"""GPU parity: the HIP relax path against the FP64 oracle, through the C ABI.

Tolerances: forces agree to 1e-12 of the largest force (bond terms are
bit-identical by construction; triple terms differ only by the last bits of
atan2); relaxed energies agree to 1e-12 relative at |F|max < 1e-10 when both
minimisers stay in one basin (small strain steps, as the reference's
protocol takes them).
"""
import sys

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(L, p, seed, le, kb, kappa=None):
    o = O.Net.lattice(L, p, seed, le).triples()
    n = N.Network.lattice(L, p, seed, le).triples()
    if kappa is not None:
        o.fiber_kappa(*kappa)
        n.fiber_kappa(*kappa)
    return o, n


@pytest.mark.parametrize("le", [True, False])
def test_force_eval_matches_oracle(le):
    _say(f"case force_eval le={le}")
    o, n = _pair(32, 0.65, 11, le, 1e-3)
    b = N.Batch([n], 1.0, 1e-3)
    b.shear(0, 0.03)
    f, e = b.eval(0)
    pos = o.pos()
    o.shear(pos, 0.03)
    fo = o.forces(1.0, 1e-3, pos)
    es, eb = o.energies(1.0, 1e-3, pos)
    scale = np.abs(fo).max()
    assert np.abs(f - fo).max() <= 1e-12 * scale
    assert abs(e[0] - es) <= 1e-12 * es
    assert abs(e[1] - eb) <= 1e-12 * eb
    np.testing.assert_array_equal(b.positions(0), pos)


def test_fire_small_matches_oracle():
    _say("case fire L=16 le")
    o, n = _pair(16, 0.7, 5, True, 1e-2)
    b = N.Batch([n], 1.0, 1e-2)
    prm = N.relax_params(dt_max=0.1, fcut=1e-10)
    b.shear(0, 0.01)
    r = b.relax(prm)[0]
    pos = o.pos()
    o.shear(pos, 0.01)
    ro = o.fire(1.0, 1e-2, pos, fp=(0.1, 0.99, 0.1, 1.1, 0.5), fcut=1e-10)
    assert r["status"] == N.CONVERGED and ro["rc"] == 0
    assert r["fmax"] < 1e-10
    et, eo = r["es"] + r["eb"], ro["es"] + ro["eb"]
    assert abs(et - eo) <= 1e-12 * eo
    np.testing.assert_allclose(b.positions(0), pos, rtol=0, atol=1e-8)


def test_batch_alone_equals_in_mix():
    _say("case batch alone vs mix")
    nets = [N.Network.lattice(L, 0.68, 100 + L, True).triples() for L in (12, 20, 16)]
    prm = N.relax_params(dt_max=0.2, fcut=1e-10)
    mix = N.Batch(nets, 1.0, 3e-3)
    mix.shear(-1, 0.02)
    rm = mix.relax(prm)
    for k, net in enumerate(nets):
        alone = N.Batch([N.Network.lattice((12, 20, 16)[k], 0.68, 100 + (12, 20, 16)[k], True).triples()],
                        1.0, 3e-3)
        alone.shear(0, 0.02)
        ra = alone.relax(prm)[0]
        assert ra["steps"] == rm[k]["steps"]
        assert ra["es"] == rm[k]["es"] and ra["eb"] == rm[k]["eb"]
        np.testing.assert_array_equal(alone.positions(0), mix.positions(k))


def test_team_kernel_matches_single_workgroup():
    _say("case team vs single workgroup")
    o, n = _pair(48, 0.7, 9, True, 1e-2)
    prm = N.relax_params(dt_max=0.2, fcut=1e-10)
    team = N.Batch([n], 1.0, 1e-2, tile_nodes=256)      # 9 tiles -> team of 9
    one = N.Batch([n], 1.0, 1e-2, tile_nodes=256, team=1)
    for b in (team, one):
        b.shear(0, 0.01)
    rt = team.relax(prm)[0]
    r1 = one.relax(prm)[0]
    assert rt["status"] == N.CONVERGED and r1["status"] == N.CONVERGED
    pos = o.pos()
    o.shear(pos, 0.01)
    ro = o.fire(1.0, 1e-2, pos, fp=(0.1, 0.99, 0.2, 1.1, 0.5), fcut=1e-10)
    eo = ro["es"] + ro["eb"]
    for r in (rt, r1):
        assert abs(r["es"] + r["eb"] - eo) <= 1e-12 * eo


def test_run_twice_bit_identical():
    _say("case determinism")
    n = N.Network.lattice(24, 0.65, 3, True).triples()
    out = []
    for _ in range(2):
        b = N.Batch([n], 1.0, 1e-3, tile_nodes=256)
        b.shear(0, 0.02)
        r = b.relax(N.relax_params(dt_max=0.2, fcut=1e-10))[0]
        out.append((r["steps"], r["es"], r["eb"], b.positions(0)))
    assert out[0][:3] == out[1][:3]
    np.testing.assert_array_equal(out[0][3], out[1][3])
