"""GPU: nr_batch_analyze_noalloc (workspace in the library's fixed device
array) against nr_batch_analyze (workspace allocated per call): the same
reductions bit for bit, on two calls in a row, the second on a smaller
network, with and without the bond strains, with the bond list staged by the
call itself and uploaded by an earlier nr_batch_analyze; and the bond limit."""
import sys

import numpy as np
import pytest

import network_relaxation_amd as N

pytestmark = pytest.mark.gpu

PRM = N.relax_params(dt_max=0.2, fcut=1e-10, max_steps=200000)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _relaxed_batch():
    nets = [N.Network.lattice(L, 0.7, 40 + L, le).triples() for L, le in ((48, True), (20, False))]
    b = N.Batch(nets, 1.0, 2e-3)
    b.shear(-1, 0.03)
    reps = b.relax(PRM)
    assert all(r["status"] == N.CONVERGED for r in reps)
    return b


def _same(a, b):
    for k in ("nonaffinity", "axial_energy", "mean_strain", "max_strain"):
        assert a[k] == b[k], (k, a[k], b[k])
    if "strains" in a:
        np.testing.assert_array_equal(a["strains"], b["strains"])


def test_noalloc_equals_allocating_entry_point():
    _say("case analyze_noalloc: staged bonds, larger then smaller network")
    b = _relaxed_batch()
    # no nr_batch_analyze call yet: the no-alloc call stages the bond lists
    first = [b.analyze(k, 0.03, strains=True, noalloc=True) for k in (0, 1)]
    without = [b.analyze(k, 0.03, noalloc=True) for k in (0, 1)]
    ref = [b.analyze(k, 0.03, strains=True) for k in (0, 1)]
    for k in (0, 1):
        _same(first[k], ref[k])
        _same(without[k], {x: ref[k][x] for x in ref[k] if x != "strains"})
    _say("case analyze_noalloc: bonds uploaded by nr_batch_analyze")
    again = [b.analyze(k, 0.03, axial=0.01, k_axial=2.0, strains=True, noalloc=True)
             for k in (0, 1)]
    ref2 = [b.analyze(k, 0.03, axial=0.01, k_axial=2.0, strains=True) for k in (0, 1)]
    for k in (0, 1):
        _same(again[k], ref2[k])


def test_noalloc_rejects_more_bonds_than_its_workspace():
    _say("case analyze_noalloc: bond limit")
    n = N.Network.lattice(600, 1.0, 1, True)
    assert n.info().n_bonds > (1 << 20)
    b = N.Batch([n], 1.0, 0.0)
    with pytest.raises(N.NRError) as e:
        b.analyze(0, 0.0, noalloc=True)
    assert e.value.code == N.E_ARG and "workspace" in str(e.value)
