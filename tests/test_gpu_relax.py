"""GPU parity: the HIP relax path against the FP64 oracle, through the C ABI.

Tolerances: forces agree to 1e-12 of the largest force (the device's 1/|r|,
1/|a|^2 and atan2 differ from the oracle's only in the last bits); relaxed
energies agree to 1e-12 relative at |F|max < 1e-10 when both
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
    mix = N.Batch(nets, 1.0, 3e-3, tiles_per_member=2)
    mix.shear(-1, 0.02)
    rm = mix.relax(prm)
    for k, net in enumerate(nets):
        alone = N.Batch([N.Network.lattice((12, 20, 16)[k], 0.68, 100 + (12, 20, 16)[k], True).triples()],
                        1.0, 3e-3, tiles_per_member=2)
        alone.shear(0, 0.02)
        ra = alone.relax(prm)[0]
        assert ra["steps"] == rm[k]["steps"]
        assert ra["es"] == rm[k]["es"] and ra["eb"] == rm[k]["eb"]
        np.testing.assert_array_equal(alone.positions(0), mix.positions(k))


def test_team_kernel_matches_single_workgroup():
    _say("case team vs single workgroup")
    o, n = _pair(48, 0.7, 9, True, 1e-2)
    prm = N.relax_params(dt_max=0.2, fcut=1e-10)
    team = N.Batch([n], 1.0, 1e-2, tile_nodes=256)      # 9 tiles -> team of 8
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


FIRE_REF = (0.1, 0.99, 0.1, 1.1, 0.5)


def test_config1_plain_fp64():
    """Config 1: 32x32 LE lattice, p=0.65, kappa=1e-3, gamma=0.01 affine start,
    relaxed to |F|max < 1e-10; energy within 1e-12 of the oracle's."""
    _say("case config1")
    o, n = _pair(32, 0.65, 1, True, 1e-3)
    b = N.Batch([n], 1.0, 1e-3)
    b.shear(0, 0.01)
    r = b.relax(N.relax_params(fcut=1e-10))[0]
    pos = o.pos()
    o.shear(pos, 0.01)
    ro = o.fire(1.0, 1e-3, pos, fp=FIRE_REF, fcut=1e-10)
    assert r["status"] == N.CONVERGED and r["fmax"] < 1e-10
    et, eo = r["es"] + r["eb"], ro["es"] + ro["eb"]
    assert abs(et - eo) <= 1e-12 * eo
    # the split moves at first order in the residual position error (stretch
    # and bend forces balance, neither vanishes), the total only at second
    # order: two minimisers stopped at |F|max < 1e-10 agree on it to ~1e-10
    assert abs(r["es"] - ro["es"]) <= 1e-9 * eo and abs(r["eb"] - ro["eb"]) <= 1e-9 * eo


def test_fiber_kappa_relax_matches_oracle():
    _say("case fiber kappa")
    o, n = _pair(24, 0.7, 21, True, 1.0, kappa=(1e-5, 1e-2, 77))
    b = N.Batch([n], 1.0, 1.0)
    b.shear(0, 0.02)
    r = b.relax(N.relax_params(dt_max=0.2, fcut=1e-10))[0]
    pos = o.pos()
    o.shear(pos, 0.02)
    ro = o.fire(1.0, 1.0, pos, fp=(0.1, 0.99, 0.2, 1.1, 0.5), fcut=1e-10)
    et, eo = r["es"] + r["eb"], ro["es"] + ro["eb"]
    assert r["status"] == N.CONVERGED
    assert abs(et - eo) <= 1e-12 * eo


def test_shear_ramp_state_carried_over():
    """Three shear increments, each relaxed from the previous state (the
    reference's ramp :966-999)."""
    _say("case ramp")
    o, n = _pair(20, 0.7, 4, True, 5e-3)
    b = N.Batch([n], 1.0, 5e-3)
    pos = o.pos()
    prm = N.relax_params(dt_max=0.2, fcut=1e-10)
    for _ in range(3):
        b.shear(0, 0.01)
        r = b.relax(prm)[0]
        o.shear(pos, 0.01)
        ro = o.fire(1.0, 5e-3, pos, fp=(0.1, 0.99, 0.2, 1.1, 0.5), fcut=1e-10)
        et, eo = r["es"] + r["eb"], ro["es"] + ro["eb"]
        assert abs(et - eo) <= 1e-12 * eo
    np.testing.assert_allclose(b.positions(0), pos, rtol=0, atol=1e-8)


def _full_lattice_energy_per_node(gamma):
    s = np.sqrt(3.0) / 2
    d60 = np.sqrt(1 + s * gamma + (s * gamma) ** 2)
    d120 = np.sqrt(1 - s * gamma + (s * gamma) ** 2)
    return 0.5 * ((d60 - 1) ** 2 + (d120 - 1) ** 2)


def test_full_lattice_closed_form_config2_size():
    """512x512 full lattice (config 2's size) under affine shear is an exact
    equilibrium: one FIRE step, energy per node from the closed form."""
    _say("case full lattice L=512")
    n = N.Network.lattice(512, 1.0, 1, True).triples()
    b = N.Batch([n], 1.0, 1e-4)
    b.shear(0, 0.05)
    r = b.relax(N.relax_params(fcut=1e-10))[0]
    assert r["steps"] == 1 and r["status"] == N.CONVERGED
    e_exp = 512 * 512 * _full_lattice_energy_per_node(0.05)
    assert abs(r["es"] - e_exp) <= 1e-11 * e_exp
    assert r["eb"] < 1e-20


def test_team_size_does_not_change_the_minimum():
    _say("case team sizes")
    n = N.Network.lattice(64, 0.7, 12, True).triples()
    prm = N.relax_params(dt_max=0.2, fcut=1e-10)
    out = []
    for tpm in (1, 4, 16):
        b = N.Batch([n], 1.0, 1e-2, tiles_per_member=tpm)
        b.shear(0, 0.01)
        r = b.relax(prm)[0]
        assert r["status"] == N.CONVERGED
        out.append(r["es"] + r["eb"])
    assert max(out) - min(out) <= 1e-12 * out[0]


def test_mixed_precision_energy_within_1e10_of_oracle():
    """FP32 stage on FP64 anchors + FP64 polish: relaxed energy within 1e-10
    relative of the FP64 oracle, |F|max below the cut."""
    _say("case mixed precision")
    for (L, p, seed, kap) in ((32, 0.65, 1, None), (32, 0.62, 8, (1e-5, 1e-2, 3))):
        o, n = _pair(L, p, seed, True, 1e-3 if kap is None else 1.0, kappa=kap)
        kb = 1e-3 if kap is None else 1.0
        b = N.Batch([n], 1.0, kb, precision=N.PREC_MIXED)
        b.shear(0, 0.01)
        r = b.relax(N.relax_params(dt_max=0.2, fcut=1e-10))[0]
        pos = o.pos()
        o.shear(pos, 0.01)
        ro = o.fire(1.0, kb, pos, fp=(0.1, 0.99, 0.2, 1.1, 0.5), fcut=1e-10)
        et, eo = r["es"] + r["eb"], ro["es"] + ro["eb"]
        assert r["status"] == N.CONVERGED and r["fmax"] < 1e-10
        assert r["fp64_steps"] < r["steps"]
        assert abs(et - eo) <= 1e-10 * eo, (et, eo)


def test_mixed_batch_alone_equals_in_mix():
    _say("case mixed alone vs mix")
    sizes = (12, 28, 18)
    mk = lambda L: N.Network.lattice(L, 0.66, 300 + L, True).triples().fiber_kappa(1e-4, 1e-2, L)
    prm = N.relax_params(dt_max=0.2, fcut=1e-9)
    mix = N.Batch([mk(L) for L in sizes], 1.0, 1.0, precision=N.PREC_MIXED, tiles_per_member=2)
    mix.shear(-1, 0.05)
    rm = mix.relax(prm)
    for k, L in enumerate(sizes):
        alone = N.Batch([mk(L)], 1.0, 1.0, precision=N.PREC_MIXED, tiles_per_member=2)
        alone.shear(0, 0.05)
        ra = alone.relax(prm)[0]
        assert (ra["steps"], ra["es"], ra["eb"]) == (rm[k]["steps"], rm[k]["es"], rm[k]["eb"])


def test_step_limit_reports_not_converged():
    _say("case step limit")
    n = N.Network.lattice(16, 0.7, 5, True).triples()
    b = N.Batch([n], 1.0, 1e-2)
    b.shear(0, 0.05)
    r = b.relax(N.relax_params(fcut=1e-12, max_steps=5))[0]
    assert r["steps"] == 5 and r["status"] == N.NOT_CONVERGED


def test_coincident_bonded_nodes_report_nonfinite():
    """A zero-length bond makes the force NaN: the relax stops with
    NONFINITE instead of running to the step limit."""
    _say("case nonfinite")
    n = N.Network.lattice(8, 1.0, 5, True).triples()
    b = N.Batch([n], 1.0, 1e-2)
    pos = n.arrays()["pos"]
    bonds = n.arrays()["bonds"]
    i, j = bonds[0]
    pos[2 * j:2 * j + 2] = pos[2 * i:2 * i + 2]
    b.set_positions(0, pos)
    r = b.relax(N.relax_params(fcut=1e-12, max_steps=1000))[0]
    assert r["status"] == N.NONFINITE and r["steps"] < 1000


@pytest.mark.parametrize("le,precision", [(True, N.PREC_FP64), (False, N.PREC_FP64),
                                          (True, N.PREC_MIXED)])
def test_shear_stress_matches_oracle_virial(le, precision):
    """sigma_xy of the relaxed state = the oracle's shear virial (dE/dgamma,
    pinned by finite differences) over the box area, to 1e-10 relative."""
    _say(f"case shear stress le={le} precision={precision}")
    o, n = _pair(24, 0.7, 31, le, 1.0, kappa=(1e-4, 1e-2, 9))
    b = N.Batch([n], 1.0, 1.0, precision=precision)
    b.shear(0, 0.05)
    r = b.relax(N.relax_params(dt_max=0.2, fcut=1e-10))[0]
    assert r["status"] == N.CONVERGED
    pos = b.positions(0)
    # the device's final image shift equals the oracle's after the same step
    p0 = o.pos()
    o.shear(p0, 0.05)
    info = n.info()
    area = info.width * (info.height if info.height > 0 else info.maxy - info.miny)
    w = o.shear_virial(1.0, 1.0, pos)
    assert abs(r["sxy"] * area - w) <= 1e-10 * abs(w), (r["sxy"] * area, w)
