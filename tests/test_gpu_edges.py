"""Edge cases of the relax path on the device, against the FP64 oracle:
bending switched off (no triples on the device), a network without bonds,
networks with bond-less nodes, and both box modes in one batch."""
import sys

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

pytestmark = pytest.mark.gpu

FIRE_REF = (0.1, 0.99, 0.1, 1.1, 0.5)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _relax_pair(L, p, seed, le, ks, kb, gamma, fcut=1e-10):
    o = O.Net.lattice(L, p, seed, le).triples()
    n = N.Network.lattice(L, p, seed, le).triples()
    b = N.Batch([n], ks, kb)
    b.shear(0, gamma)
    r = b.relax(N.relax_params(fcut=fcut))[0]
    pos = o.pos()
    o.shear(pos, gamma)
    ro = o.fire(ks, kb, pos, fp=FIRE_REF, fcut=fcut)
    return b, r, o, pos, ro


@pytest.mark.parametrize("le", [True, False])
def test_bending_off_matches_oracle(le):
    """kb = 0: the reference skips the bending kernels; the device drops the
    triples from its tiles.  Same minimum as the oracle's stretch-only FIRE."""
    _say(f"case kb=0 le={le}")
    b, r, o, pos, ro = _relax_pair(20, 0.8, 4, le, 1.0, 0.0, 0.01)
    assert r["status"] == N.CONVERGED and ro["rc"] == 0
    assert r["eb"] == 0.0 and ro["eb"] == 0.0
    assert r["es"] == pytest.approx(ro["es"], rel=1e-12)


def test_network_without_bonds():
    """p = 0: every node is bond-less, the force is zero at the start; the
    relax stops at once and the positions are the affine start."""
    _say("case no bonds")
    b, r, o, pos, ro = _relax_pair(12, 0.0, 1, True, 1.0, 1e-3, 0.05)
    assert o.e == 0
    assert r["status"] == N.CONVERGED and r["es"] == 0.0 and r["eb"] == 0.0
    assert r["steps"] == ro["steps"]
    np.testing.assert_array_equal(b.positions(0), pos)


@pytest.mark.parametrize("le", [True, False])
def test_bondless_nodes_keep_their_affine_position(le):
    """A dilute network (p = 0.35) has bond-less nodes: they feel no force
    and stay where the affine step put them, while the rest relaxes to the
    oracle's minimum."""
    _say(f"case dilute le={le}")
    b, r, o, pos, ro = _relax_pair(20, 0.35, 8, le, 1.0, 1e-2, 0.02)
    bonds = o.bonds()
    deg = np.bincount(bonds.reshape(-1), minlength=o.n)
    iso = np.flatnonzero(deg == 0)
    assert len(iso) > 0
    assert r["status"] == N.CONVERGED and ro["rc"] == 0
    x = b.positions(0).reshape(-1, 2)
    np.testing.assert_array_equal(x[iso], pos.reshape(-1, 2)[iso])
    e, eo = r["es"] + r["eb"], ro["es"] + ro["eb"]
    assert e == pytest.approx(eo, rel=1e-10, abs=1e-20)


def test_both_box_modes_in_one_batch():
    """A Lees-Edwards and a clamped-row network relaxed in one batch equal
    each relaxed alone (fixed tiles per member), bit for bit."""
    _say("case box modes in one batch")
    mk = [lambda: N.Network.lattice(24, 0.7, 3, True).triples(),
          lambda: N.Network.lattice(24, 0.7, 3, False).triples()]
    prm = N.relax_params(dt_max=0.2, fcut=1e-10)
    mix = N.Batch([f() for f in mk], 1.0, 2e-3, tiles_per_member=2)
    mix.shear(-1, 0.03)
    rm = mix.relax(prm)
    for k, f in enumerate(mk):
        alone = N.Batch([f()], 1.0, 2e-3, tiles_per_member=2)
        alone.shear(0, 0.03)
        ra = alone.relax(prm)[0]
        assert (ra["steps"], ra["es"], ra["eb"]) == (rm[k]["steps"], rm[k]["es"], rm[k]["eb"])
        np.testing.assert_array_equal(alone.positions(0), mix.positions(k))
