"""The oracle's post-relax reductions (oracle/analysis.c, restating
src/na_calc.cpp:308-326, src/strain_calc.cpp:96-120 and
src/axial_contribution.cpp:173-190), pinned by closed forms: a full triangular
lattice under affine shear gamma has no non-affine displacement, bond strains
sqrt(1 +- (sqrt(3)/2) gamma + (3/4) gamma^2) - 1 on the slanted bonds and 0 on
the horizontal ones, and axial energy k/2 (gamma sqrt(3)/4)^2 per slanted bond.
"""
import math

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("le", [True, False])
def test_affine_shear_of_full_lattice(le):
    L, g, k = 8, 0.03, 1.7
    net = O.Net.lattice(L, 1.0, 1, le_mode=le).triples()
    x0 = net.pos()
    x = x0.copy()
    net.shear(x, g)
    shift = g * net.s.height if le else 0.0
    r = net.analyze(x0, x, shift, g, 0.0, k)
    assert r["nonaffinity"] < 1e-28
    d = x0.reshape(-1, 2)
    b = net.bonds()
    w = net.bond_wraps()
    dy = d[b[:, 1], 1] + w[:, 1] * net.s.height - d[b[:, 0], 1]
    dx = d[b[:, 1], 0] + w[:, 0] * net.s.width - d[b[:, 0], 0]
    expect = np.where(np.abs(dy) < 1e-9, 0.0,
                      np.sqrt(1.0 + np.sign(dx * dy) * math.sqrt(3.0) / 2.0 * g + 0.75 * g * g) - 1.0)
    assert np.abs(r["strains"] - expect).max() < 1e-14
    slanted = int((np.abs(dy) > 1e-9).sum())
    assert slanted == (2 * L * L if le else 2 * L * (L - 1))
    assert r["axial_energy"] == pytest.approx(k * 0.5 * (g * math.sqrt(3.0) / 4.0) ** 2 * slanted,
                                              rel=1e-12)
    assert r["max_strain"] == pytest.approx(expect.max(), rel=1e-14)
    assert r["mean_strain"] == pytest.approx(expect.mean(), rel=1e-12)


def test_nonaffinity_of_a_uniform_translation():
    # a rigid translation t against a zero affine field: |t|^2 at every node
    net = O.Net.lattice(6, 0.7, 2, le_mode=False).triples()
    x0 = net.pos()
    x = x0 + np.tile([0.3, -0.4], net.n)
    r = net.analyze(x0, x, 0.0, 0.0, 0.0)
    assert r["nonaffinity"] == pytest.approx(0.25, rel=1e-14)
    assert r["axial_energy"] < 1e-28 and r["max_strain"] < 1e-15
