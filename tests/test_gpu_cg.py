"""GPU nonlinear CG (Fletcher-Reeves and Polak-Ribiere+, secant line search)
through the C ABI.

The reference only names the method (large_strain_types.hpp:31-34); the FP64
oracle's CG is the algorithm both follow.  The device forms f.d and |d|^2 of
each new direction from the previous trial's reductions instead of summing
them again, so trajectories agree to rounding, not bit for bit: relaxed
energies must agree with the oracle's CG and with the FIRE minimum to 1e-12
relative at |F|max < 1e-10 (one basin, small strain), and the iteration
counts stay close.
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


@pytest.mark.parametrize("L,le,pr", [(16, True, False), (24, False, False), (64, True, False),
                                     (16, True, True), (24, False, True), (64, True, True)])
def test_cg_matches_oracle_cg_and_fire_minimum(L, le, pr):
    """Fletcher-Reeves (pr False) and Polak-Ribiere+ (pr True)."""
    _say(f"case cg L={L} le={le} pr={pr}")
    o = O.Net.lattice(L, 0.7, 40 + L, le).triples()
    n = N.Network.lattice(L, 0.7, 40 + L, le).triples()
    kb = 4e-3
    cg = N.relax_params(fcut=1e-10, method=N.METHOD_CG_PR if pr else N.METHOD_CG)
    b = N.Batch([n], 1.0, kb)
    b.shear(0, 0.02)
    r = b.relax(cg)[0]
    assert r["status"] == N.CONVERGED and r["fmax"] < 1e-10
    pos = o.pos()
    o.shear(pos, 0.02)
    pos_cg = pos.copy()
    ro = o.cg(1.0, kb, pos_cg, fcut=1e-10, pr=pr)
    assert ro["rc"] == 0
    et, eo = r["es"] + r["eb"], ro["es"] + ro["eb"]
    assert abs(et - eo) <= 1e-12 * eo, (et, eo)
    assert abs(r["steps"] - ro["steps"]) <= max(5, 0.25 * ro["steps"]), (r["steps"], ro["steps"])
    # FIRE reaches the same minimum
    bf = N.Batch([n], 1.0, kb)
    bf.shear(0, 0.02)
    rf = bf.relax(N.relax_params(dt_max=0.2, fcut=1e-10))[0]
    ef = rf["es"] + rf["eb"]
    assert abs(et - ef) <= 1e-12 * ef, (et, ef)
    # (positions are compared only through the energy: the minimum is defined
    # up to rigid translations and floppy modes, which the methods leave apart)


def test_cg_ramp_equals_sequential_and_step_limit():
    _say("case cg ramp")
    n = N.Network.lattice(32, 0.68, 7, True).triples().fiber_kappa(1e-4, 1e-2, 7)
    cg = N.relax_params(fcut=1e-10, method=N.METHOD_CG)
    yr = n.info().height
    incs = np.array([(0.01 * yr, 0.0)] * 3)
    rb = N.Batch([n], 1.0, 1.0)
    ramp = rb.ramp(incs, cg)[0]
    sb = N.Batch([n], 1.0, 1.0)
    for i in range(3):
        sb.strain(0, incs[i][0], 0.0)
        r = sb.relax(cg)[0]
        assert (r["steps"], r["es"], r["eb"]) == (ramp[i]["steps"], ramp[i]["es"], ramp[i]["eb"])
    np.testing.assert_array_equal(rb.positions(0), sb.positions(0))
    lim = N.Batch([n], 1.0, 1.0)
    lim.shear(0, 0.05)
    r = lim.relax(N.relax_params(fcut=1e-12, max_steps=3, method=N.METHOD_CG))[0]
    assert r["steps"] == 3 and r["status"] == N.NOT_CONVERGED


def test_cg_team_relax_repeats_bit_for_bit():
    """CG on a team of four workgroups (one tile each) relaxed three times:
    identical steps, energies and positions.  Team members exchange state
    through 16-byte stores; a store whose data registers are overwritten
    too early would show here as a run-to-run difference."""
    _say("case cg team repeat")
    n = N.Network.lattice(32, 0.68, 7, True).triples().fiber_kappa(1e-4, 1e-2, 7)
    out = []
    for _ in range(3):
        b = N.Batch([n], 1.0, 1.0, tiles_per_member=1)
        b.shear(0, 0.01)
        r = b.relax(N.relax_params(fcut=1e-10, method=N.METHOD_CG))[0]
        assert r["status"] == N.CONVERGED
        out.append(((r["steps"], r["es"], r["eb"]), b.positions(0)))
    for key, pos in out[1:]:
        assert key == out[0][0]
        np.testing.assert_array_equal(pos, out[0][1])


def test_one_method_per_relax_call():
    _say("case method mix rejected")
    nets = [N.Network.lattice(12, 0.7, s, True).triples() for s in (1, 2)]
    b = N.Batch(nets, 1.0, 1e-3)
    with pytest.raises(N.NRError):
        b.relax([N.relax_params(), N.relax_params(method=N.METHOD_CG)])
