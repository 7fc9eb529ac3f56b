"""Device post-relax reductions (nr_batch_analyze) against the oracle's
restatement of na_calc / strain_calc / axial_contribution, on the relaxed
positions the device holds; and the full-lattice closed form on the device."""
import math

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

pytestmark = pytest.mark.gpu

FIRE = dict(alpha0=0.1, f_alpha=0.99, dt_max=0.1, f_inc=1.1, f_dec=0.5, nmin=5, mass=1.0)


def test_device_reductions_match_oracle():
    specs = [(24, 0.7, 3, True, 0.1), (20, 0.75, 4, False, 0.05), (18, 0.68, 9, True, 0.3)]
    nets, onets = [], []
    for L, p, seed, le, g in specs:
        n = N.Network.lattice(L, p, seed, le).triples()
        n.fiber_kappa(1e-4, 1e-2, seed)
        o = O.Net.lattice(L, p, seed, le).triples()
        o.fiber_kappa(1e-4, 1e-2, seed)
        nets.append(n)
        onets.append(o)
    print("analysis batch:", specs, flush=True)
    b = N.Batch(nets, 1.0, 1.0)
    for k, s in enumerate(specs):
        b.shear(k, s[4])
    reps = b.relax(N.relax_params(fcut=1e-10, criterion=N.CRIT_MAX, **FIRE))
    assert all(r["status"] == N.CONVERGED for r in reps)
    for k, (L, p, seed, le, g) in enumerate(specs):
        o = onets[k]
        x = b.positions(k)
        shift = g * o.s.height if le else 0.0
        dev = b.analyze(k, g, 0.0, 2.0, strains=True)
        ref = o.analyze(o.pos(), x, shift, g, 0.0, 2.0)
        assert dev["nonaffinity"] == pytest.approx(ref["nonaffinity"], rel=1e-12)
        assert dev["axial_energy"] == pytest.approx(ref["axial_energy"], rel=1e-12)
        assert dev["mean_strain"] == pytest.approx(ref["mean_strain"], rel=1e-11, abs=1e-16)
        assert dev["max_strain"] == pytest.approx(ref["max_strain"], rel=1e-13)
        assert np.abs(dev["strains"] - ref["strains"]).max() < 1e-15
        assert dev["nonaffinity"] > 0.0


def test_device_reductions_full_lattice_closed_form():
    L, g = 16, 0.02
    net = N.Network.lattice(L, 1.0, 1, True).triples()
    print("full lattice L=16 gamma=0.02", flush=True)
    b = N.Batch([net], 1.0, 1e-3)
    b.shear(0, g)
    r = b.relax(N.relax_params(fcut=1e-10, criterion=N.CRIT_MAX, **FIRE))[0]
    assert r["status"] == N.CONVERGED
    a = b.analyze(0, g, strains=True)
    assert a["nonaffinity"] < 1e-26
    slanted = 2 * L * L
    hi = math.sqrt(1.0 + math.sqrt(3.0) / 2.0 * g + 0.75 * g * g) - 1.0
    assert a["max_strain"] == pytest.approx(hi, rel=1e-12)
    assert a["axial_energy"] == pytest.approx(0.5 * (g * math.sqrt(3.0) / 4.0) ** 2 * slanted,
                                              rel=1e-10)
    assert (np.abs(a["strains"]) < 1e-12).sum() == L * L  # horizontal bonds
