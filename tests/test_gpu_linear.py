"""Device small-strain linear solve (nr_linear_response, batched:
nr_linear_response_batch) against the oracle's
restatement of the reference's sparse_relax.cpp (itself pinned by
tests/test_linear.py): energies, residual and the minimum-norm displacement
field; and the full-lattice closed form on the device."""
import math

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,p,seed,le,kappa,mu", [
    (24, 0.8, 3, True, None, 0.0),
    (20, 0.75, 4, False, (1e-4, 1e-2, 5), 0.0),
    (16, 0.7, 9, True, (1e-3, 1e-1, 2), 0.05),
])
def test_device_linear_matches_oracle(L, p, seed, le, kappa, mu):
    print(f"linear L={L} p={p} le={le} mu={mu}", flush=True)
    n = N.Network.lattice(L, p, seed, le).triples()
    o = O.Net.lattice(L, p, seed, le).triples()
    kb = 1e-2
    if kappa:
        n.fiber_kappa(*kappa)
        o.fiber_kappa(*kappa)
        kb = 0.0
    g = 0.01
    dev = N.linear_response(n, 1.0, kb, g, mu=mu, displacements=True)
    ref = o.linear_response(1.0, kb, g, mu=mu)
    assert dev["status"] == N.CONVERGED and ref["rc"] == 0
    e_d, e_r = dev["es"] + dev["eb"], ref["es"] + ref["eb"]
    assert e_d == pytest.approx(e_r, rel=1e-9)
    assert dev["es"] == pytest.approx(ref["es"], rel=1e-8)
    assert dev["eb"] == pytest.approx(ref["eb"], rel=1e-8, abs=1e-14)
    assert dev["residual"] < 1e-10
    scale = np.abs(ref["u"]).max()
    assert np.abs(dev["u"] - ref["u"]).max() < 1e-6 * scale
    assert abs(dev["iters"] - ref["iters"]) <= max(3, ref["iters"] // 50)


def test_device_linear_full_lattice_closed_form():
    L, g = 16, 0.02
    n = N.Network.lattice(L, 1.0, 1, True).triples()
    r = N.linear_response(n, 1.0, 1e-2, g, displacements=True)
    assert r["status"] == N.CONVERGED and r["iters"] == 0
    assert r["es"] == pytest.approx(0.5 * (g * math.sqrt(3.0) / 4.0) ** 2 * 2 * L * L, rel=1e-12)
    assert r["eb"] < 1e-25


def test_device_linear_repeats_bit_for_bit():
    n = N.Network.lattice(24, 0.7, 6, False).triples().fiber_kappa(1e-4, 1e-2, 3)
    a = N.linear_response(n, 1.0, 0.0, 0.01, displacements=True)
    b = N.linear_response(n, 1.0, 0.0, 0.01, displacements=True)
    assert (a["iters"], a["es"], a["eb"]) == (b["iters"], b["es"], b["eb"])
    np.testing.assert_array_equal(a["u"], b["u"])


def test_batched_solves_equal_single_solves():
    """One device batch of systems (networks of different sizes and box
    modes, load steps sharing a network's matrix, a system that starts
    converged, one stopped by its iteration cap) gives every system the
    results of its solve alone, bit for bit."""
    le = N.Network.lattice(24, 0.7, 3, True).triples().fiber_kappa(1e-4, 1e-2, 3)
    box = N.Network.lattice(20, 0.75, 4, False).triples()
    big = N.Network.lattice(40, 0.65, 5, True).triples()
    full = N.Network.lattice(12, 1.0, 1, True).triples()
    jobs = [dict(net=le, ks=1.0, kb=0.0, shear=0.01),
            dict(net=box, ks=1.0, kb=1e-2, shear=0.02, compression=0.01, affine=False),
            dict(net=le, ks=1.0, kb=0.0, shear=0.03, compression=0.005),
            dict(net=big, ks=1.0, kb=1e-3, shear=0.01, mu=0.02),
            dict(net=full, ks=1.0, kb=1e-2, shear=0.02),
            dict(net=big, ks=1.0, kb=1e-3, shear=0.01, max_iter=40)]
    print(f"linear batch of {len(jobs)}", flush=True)
    batch = N.linear_response_batch(jobs, displacements=True)
    assert batch[4]["iters"] == 0 and batch[5]["iters"] == 40
    assert batch[5]["status"] == N.NOT_CONVERGED
    for j, b in zip(jobs, batch):
        a = N.linear_response(j["net"], j["ks"], j["kb"], j["shear"],
                              compression=j.get("compression", 0.0), mu=j.get("mu", 0.0),
                              affine=j.get("affine", True), max_iter=j.get("max_iter", 1000000),
                              displacements=True)
        assert (a["iters"], a["status"], a["es"], a["eb"], a["residual"]) == \
            (b["iters"], b["status"], b["es"], b["eb"], b["residual"])
        np.testing.assert_array_equal(a["u"], b["u"])


def test_concurrent_solves_equal_sequential():
    """Solves from several host threads at once (each on its own device
    stream) give the results of the same solves one at a time, bit for
    bit."""
    from concurrent.futures import ThreadPoolExecutor
    nets = [N.Network.lattice(L, 0.7, s, True).triples().fiber_kappa(1e-4, 1e-2, s)
            for L, s in ((24, 3), (32, 4), (20, 5), (28, 6))]

    def solve(n):
        return N.linear_response(n, 1.0, 0.0, 0.01, displacements=True)

    seq = [solve(n) for n in nets]
    with ThreadPoolExecutor(len(nets)) as ex:
        conc = list(ex.map(solve, nets))
    for a, b in zip(seq, conc):
        assert (a["iters"], a["es"], a["eb"], a["status"]) == (b["iters"], b["es"], b["eb"], b["status"])
        np.testing.assert_array_equal(a["u"], b["u"])
