"""GPU: a job whose forces turn non-finite stops with NR_NONFINITE and leaves
the networks it shares a launch with untouched -- in a batch (one persistent
launch) and in the serving loop's mix -- bit for bit equal to the same
networks relaxed alone.  The poison is a bond of zero length (two bonded
nodes moved onto each other): its force ks (1 - L/|r|) r is 0 x inf = NaN on
the first pass, as the reference's axial_forces (large_strain_relax.cu:
401-404) would compute it."""
import sys

import numpy as np
import pytest

import network_relaxation_amd as N

pytestmark = pytest.mark.gpu

SIZES = (16, 24, 20)
PRM = N.relax_params(dt_max=0.2, fcut=1e-10, max_steps=200000)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _net(L):
    return N.Network.lattice(L, 0.68, 300 + L, True).triples()


def _collapse_first_bond(pos, net):
    i, j = net.arrays()["bonds"][0]
    pos = pos.copy()
    pos[2 * j:2 * j + 2] = pos[2 * i:2 * i + 2]
    return pos


@pytest.mark.parametrize("precision", [N.PREC_FP64, N.PREC_MIXED])
def test_poisoned_network_in_a_batch(precision):
    _say(f"case poisoned batch precision={precision}")
    nets = [_net(L) for L in SIZES]
    mix = N.Batch(nets, 1.0, 3e-3, precision=precision, tiles_per_member=2)
    mix.shear(-1, 0.02)
    mix.set_positions(1, _collapse_first_bond(nets[1].arrays()["pos"], nets[1]))
    rm = mix.relax(PRM)
    assert rm[1]["status"] == N.NONFINITE and rm[1]["steps"] < PRM.max_steps
    for k in (0, 2):
        alone = N.Batch([_net(SIZES[k])], 1.0, 3e-3, precision=precision, tiles_per_member=2)
        alone.shear(0, 0.02)
        ra = alone.relax(PRM)[0]
        assert rm[k]["status"] == N.CONVERGED
        assert (ra["steps"], ra["es"], ra["eb"]) == (rm[k]["steps"], rm[k]["es"], rm[k]["eb"])
        np.testing.assert_array_equal(alone.positions(0), mix.positions(k))


def _submit_all(srv, nets, poisoned):
    ids = {}
    for k, n in enumerate(nets):
        h = n.info().height
        incs = np.array([(0.02 * h, 0.0), (0.01 * h, 0.0)])
        ids[srv.submit(n, 1.0, 3e-3, PRM, incs, keep_positions=True)] = k
    res = {}
    while len(res) < len(ids):
        srv.step()
        for r in srv.results():
            res[ids[r["job"]]] = r
    return res


def test_poisoned_job_in_the_server():
    _say("case poisoned server job")
    nets = [_net(L) for L in SIZES]
    a = nets[1].arrays()
    info = nets[1].info()
    bad = N.Network.from_arrays(_collapse_first_bond(a["pos"], nets[1]), a["bonds"],
                                a["bond_wraps"], info.width, info.height).triples()
    srv = N.Server(precision=N.PREC_MIXED, slice_steps=500)
    res = _submit_all(srv, [nets[0], bad, nets[2]], poisoned=1)
    assert res[1]["status"] == N.NONFINITE
    for k in (0, 2):
        alone = _submit_all(N.Server(precision=N.PREC_MIXED, slice_steps=500), [_net(SIZES[k])],
                            poisoned=None)[0]
        assert res[k]["status"] == N.CONVERGED
        assert [(x["steps"], x["es"], x["eb"]) for x in alone["reports"]] == \
               [(x["steps"], x["es"], x["eb"]) for x in res[k]["reports"]]
        np.testing.assert_array_equal(alone["pos"], res[k]["pos"])
