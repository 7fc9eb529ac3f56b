"""The oracle's small-strain linear response (oracle/linear.c, restating
src/sparse_relax.cpp): stiffness assembly, the reduced solve, energies.

Pinned by what the mathematics fixes, not by retyping the assembly:
  * the bending and (unit-length) stretching blocks equal the Hessian of the
    reference's linearised energies (get_stretch_shear_energy :49-70,
    get_bending_energy :73-96), taken by polarisation of those energies;
  * the reference's own forms where they differ from that Hessian: ks d d^T/l^3
    for a bond of length l != 1 (:286-322) and the gel block without x-y
    coupling (:390-428), on single bonds;
  * a full lattice under affine shear is an exact equilibrium (every free node
    is an inversion centre): no correction, E = ks (gamma sqrt(3)/4)^2 per
    slanted bond pair, no bending;
  * the reduced system's residual;
  * agreement with the nonlinear FIRE minimum to O(gamma): the relative energy
    difference shrinks tenfold when the strain does;
  * the minimum-norm choice: bond-less nodes do not move.
"""
import math

import numpy as np
import pytest

import oracle as O


def _hessian_by_polarisation(net, ks, kb, mu=0.0):
    """H_ij of E(u) = u^T H u / 2, from the linearised energies alone."""
    n2 = 2 * net.n
    E = np.zeros(n2)
    for i in range(n2):
        u = np.zeros(n2)
        u[i] = 1.0
        es, eb = net.linear_energies(ks, kb, u, 0.0, 0.0, mu)
        E[i] = es + eb
    H = np.diag(2.0 * E)
    for i in range(n2):
        for j in range(i + 1, n2):
            u = np.zeros(n2)
            u[i] = u[j] = 1.0
            es, eb = net.linear_energies(ks, kb, u, 0.0, 0.0, mu)
            H[i, j] = H[j, i] = es + eb - E[i] - E[j]
    return H


@pytest.mark.parametrize("le", [True, False])
def test_stiffness_is_hessian_of_linearised_energy(le):
    net = O.Net.lattice(4, 0.8, 11, le_mode=le).triples()
    net.fiber_kappa(1e-3, 1e-1, 5)
    K = net.linear_stiffness(1.3, 0.0)  # per-fiber kappa replaces kb
    H = _hessian_by_polarisation(net, 1.3, 0.0)
    assert net.t > 0
    scale = np.abs(H).max()
    assert np.abs(K - K.T).max() == 0.0
    assert np.abs(K - H).max() < 1e-12 * scale
    assert np.linalg.eigvalsh(K).min() > -1e-12 * scale


def test_stretch_block_uses_inverse_cube_of_length():
    # bonds of length 2 along x and sqrt(2) at 45 degrees
    pos = [[0.0, 0.0], [2.0, 0.0], [1.0, 1.0], [0.0, 2.0]]
    net = O.Net.from_arrays(pos, [[0, 1], [1, 3]], [[0, 0], [0, 0]], 10.0)
    ks = 1.7
    K = net.linear_stiffness(ks, 0.0)
    assert K[0, 0] == pytest.approx(ks * 4.0 / 8.0, rel=1e-15)
    assert K[0, 2] == pytest.approx(-ks * 4.0 / 8.0, rel=1e-15)
    assert K[1, 1] == 0.0
    # bond 1 -> 3: d = (-2, 2), l = 2 sqrt(2)
    l3 = (2.0 * math.sqrt(2.0)) ** 3
    assert K[2, 3] == pytest.approx(ks * (-2.0) * 2.0 / l3, rel=1e-14)
    assert K[2, 7] == pytest.approx(-ks * (-2.0) * 2.0 / l3, rel=1e-14)


def test_gel_block_has_no_xy_coupling():
    pos = [[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]
    net = O.Net.from_arrays(pos, [[0, 1]], [[0, 0]], 10.0)
    mu = 0.3
    K = net.linear_stiffness(0.0, 0.0, mu)
    assert K[0, 0] == 0.0 and K[0, 2] == 0.0        # mu (1 - 1)
    assert K[1, 1] == pytest.approx(mu) and K[1, 3] == pytest.approx(-mu)
    assert K[0, 1] == 0.0 and K[0, 3] == 0.0 and K[1, 2] == 0.0


@pytest.mark.parametrize("le", [True, False])
def test_full_lattice_affine_shear_closed_form(le):
    L, g, ks = 8, 0.01, 1.0
    net = O.Net.lattice(L, 1.0, 1, le_mode=le).triples()
    r = net.linear_response(ks, 1e-2, g)
    assert r["rc"] == 0
    per_bond = 0.5 * ks * (g * math.sqrt(3.0) / 4.0) ** 2  # slanted bonds; 0 for 0 degrees
    slanted = 2 * L * L if le else 2 * L * (L - 1)
    assert r["es"] == pytest.approx(per_bond * slanted, rel=1e-12)
    assert r["eb"] < 1e-25
    # no correction: u is the affine field itself
    pos = net.pos().reshape(-1, 2)
    miny = pos[:, 1].min()
    assert np.abs(r["u"][0::2] - g * (pos[:, 1] - miny)).max() < 1e-14
    assert np.abs(r["u"][1::2]).max() == 0.0


@pytest.mark.parametrize("le", [True, False])
def test_reduced_system_residual(le):
    net = O.Net.lattice(24, 0.7, 3, le_mode=le).triples()
    net.fiber_kappa(1e-4, 1e-2, 9)
    r = net.linear_response(1.0, 0.0, 0.02, mu=0.05)
    assert r["rc"] == 0
    assert r["residual"] < 1e-10
    assert r["es"] > 0 and r["eb"] > 0


@pytest.mark.parametrize("le", [True, False])
def test_agrees_with_fire_to_first_order_in_strain(le):
    rel = []
    for g in (1e-3, 1e-2):
        net = O.Net.lattice(16, 0.8, 5, le_mode=le).triples()
        lin = net.linear_response(1.0, 1e-2, g)
        pos = net.pos()
        net.shear(pos, g)
        fire = net.fire(1.0, 1e-2, pos, fcut=1e-13)
        assert fire["rc"] == 0
        e_f = fire["es"] + fire["eb"]
        rel.append(abs(lin["es"] + lin["eb"] - e_f) / e_f)
    assert rel[0] < 2e-4
    assert 5.0 < rel[1] / rel[0] < 20.0


def test_bondless_nodes_do_not_move():
    net = O.Net.lattice(16, 0.3, 2, le_mode=False).triples()
    bonds = net.bonds().reshape(-1, 2)
    deg = np.bincount(bonds.reshape(-1), minlength=net.n)
    assert (deg == 0).any()
    g = 0.01
    r = net.linear_response(1.0, 1e-3, g)
    pos = net.pos().reshape(-1, 2)
    bc = g * (pos[:, 1] - pos[:, 1].min())
    iso = np.where(deg == 0)[0]
    assert np.array_equal(r["u"][2 * iso], bc[iso])
    assert np.array_equal(r["u"][2 * iso + 1], np.zeros(len(iso)))


def test_top_row_displacement_mode():
    net = O.Net.lattice(12, 0.75, 4, le_mode=False).triples()
    s, c = 0.05, 0.02
    r = net.linear_response(1.0, 1e-3, s, c, affine=False)
    assert r["rc"] == 0 and r["residual"] < 1e-10
    pos = net.pos().reshape(-1, 2)
    top = pos[:, 1] > pos[:, 1].max() - 1e-10
    bot = pos[:, 1] < pos[:, 1].min() + 1e-10
    assert np.array_equal(r["u"][0::2][top], np.full(top.sum(), s))
    assert np.array_equal(r["u"][1::2][top], np.full(top.sum(), -c))
    assert not r["u"][0::2][bot].any() and not r["u"][1::2][bot].any()


def test_periodic_box_rejects_top_row_mode():
    net = O.Net.lattice(8, 0.8, 4, le_mode=True).triples()
    assert net.linear_response(1.0, 1e-3, 0.01, affine=False)["rc"] == 7  # OR_E_ARG
