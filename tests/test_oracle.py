"""CPU checks that pin the FP64 oracle (no GPU).

The reference ships no tests, fixtures or golden outputs, so each oracle
function is pinned by something independent of its own code: hand-enumerated
fixtures, closed forms, finite differences, invariants, or a hand-derived
trajectory.  A planted mistake (sign, index, dropped term, wrong update order)
fails at least one of these.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- read_network (large_strain_relax.cu:42-114) --------------------------

def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_read_network_rest_lengths_and_wraps(tmp_path):
    path = _write(tmp_path, "n.dat", "3 2 4.0\n0 0\n3.5 0.5\n1 0\n1 0 -1\n0 2 0\n")
    n = O.Net.read(path)
    assert (n.n, n.e) == (3, 2)
    # bond (1, 0, -1): dx = x0 + 4*(-1) - x1 = -7.5, dy = -0.5
    np.testing.assert_array_equal(n.rest(), [math.sqrt(7.5 ** 2 + 0.25), 1.0])
    np.testing.assert_array_equal(n.bond_wraps()[:, 0], [-1, 0])


def test_read_network_skips_malformed_lines(tmp_path):
    path = _write(tmp_path, "n.dat", "2 1 5\njunk\n0 0\n\n1 1\nnot an edge\n0 1 0\n")
    n = O.Net.read(path)
    assert (n.n, n.e) == (2, 1)
    np.testing.assert_array_equal(n.pos(), [0, 0, 1, 1])


@pytest.mark.parametrize("text,code", [
    ("2 1\n0 0\n1 1\n0 1 0\n", 2),            # header needs 3 numbers
    ("0 1 5\n", 2),                           # N <= 0
    ("2 1 -1\n0 0\n1 1\n0 1 0\n", 2),         # negative width
    ("3 1 5\n0 0\n1 1\n", 3),                 # too few points
    ("2 1 5\n0 0\n1 1", 3),                   # last point line without newline: eof
    ("2 1 5\n0 0\n1 1\n0 2 0\n", 4),          # index out of range
    ("2 2 5\n0 0\n1 1\n0 1 0\n", 5),          # too few edges
])
def test_read_network_rejections(tmp_path, text, code):
    path = _write(tmp_path, "bad.dat", text)
    with pytest.raises(ValueError, match=f"code {code}"):
        O.Net.read(path)


def test_clamp_codes_reference_mode(tmp_path):
    # y within 1e-10 of min or max y is frozen (:896-902)
    path = _write(tmp_path, "c.dat", "4 0 10\n0 0\n1 5e-11\n2 1\n3 2\n")
    n = O.Net.read(path)
    np.testing.assert_array_equal(n.fcode(), [0, 0, 3, 0])


def test_lees_edwards_header_and_y_wrap(tmp_path):
    path = _write(tmp_path, "le.dat", "2 1 4 3\n0 2.5\n0 0.5\n0 1 0 1\n")
    n = O.Net.read(path)
    # dy = 0.5 + 3*1 - 2.5 = 1
    np.testing.assert_array_equal(n.rest(), [1.0])
    np.testing.assert_array_equal(n.fcode(), [3, 3])


# ---- get_triples (:116-184) -------------------------------------------------

def _golden_triples():
    with open(os.path.join(GOLDEN, "triples_small.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("case", _golden_triples(), ids=lambda c: c["name"])
def test_triples_hand_enumerated(case):
    pos = np.array(case["pos"], float)
    bonds = np.array([b[:2] for b in case["bonds"]], np.int32)
    wraps = np.array([[b[2], 0] for b in case["bonds"]], np.int8)
    n = O.Net.from_arrays(pos, bonds, wraps, case["width"]).triples()
    got = [list(t) + [int(w[0]), int(w[2])] for t, w in zip(n.tris(), n.tri_wraps())]
    assert got == case["triples"]


def _brute_triples(pos, bonds, wraps, W, H):
    """All antiparallel neighbour pairs, hinge ascending, pairs in the
    neighbour list order built from the edge list."""
    adj = {}
    for b, (i, j) in enumerate(bonds):
        adj.setdefault(i, []).append((j, wraps[b][0], wraps[b][1]))
        adj.setdefault(j, []).append((i, -wraps[b][0], -wraps[b][1]))
    out = []
    for h in sorted(adj):
        nb = adj[h]
        for a in range(len(nb)):
            for c in range(a + 1, len(nb)):
                (n1, x1, y1), (n2, x2, y2) = nb[a], nb[c]
                u = np.array([pos[n1][0] + W * x1 - pos[h][0], pos[n1][1] + H * y1 - pos[h][1]])
                v = np.array([pos[n2][0] + W * x2 - pos[h][0], pos[n2][1] + H * y2 - pos[h][1]])
                cosang = u @ v / (np.linalg.norm(u) * np.linalg.norm(v))
                if cosang < -1 + 1e-9:
                    out.append((h, n1, n2, x1, y1, x2, y2))
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("le", [False, True])
def test_triples_brute_force_small_lattices(seed, le):
    n = O.Net.lattice(6, 0.7, seed, le).triples()
    pos = n.pos().reshape(-1, 2)
    got = [tuple(int(v) for v in t) + tuple(int(w) for w in tw)
           for t, tw in zip(n.tris(), n.tri_wraps())]
    got = [(g[0], g[1], g[2], g[3], g[4], g[5], g[6]) for g in got]
    exp = _brute_triples(pos, n.bonds(), n.bond_wraps(), n.s.width, n.s.height)
    assert got == exp


def test_full_lattice_triple_count():
    # every node of a full periodic triangular lattice is the hinge of 3 lines
    n = O.Net.lattice(8, 1.0, 1, True).triples()
    assert n.t == 3 * n.n and n.e == 3 * n.n


# ---- forces and energies (:386-523) -----------------------------------------

def _energy(n, ks, kb, pos):
    es, eb = n.energies(ks, kb, pos)
    return es + eb


@pytest.mark.parametrize("le", [True, False])
def test_forces_are_minus_energy_gradient(le):
    n = O.Net.lattice(8, 0.8, 4, le).triples()
    n.fiber_kappa(1e-3, 1e-1, 9)
    rng = np.random.RandomState(0)
    pos = n.pos() + 0.05 * rng.randn(2 * n.n)
    n.shear(pos, 0.05)
    f = n.forces(1.0, 1.0, pos)
    free = np.repeat(n.fcode() == 3, 2)
    h = 1e-6
    for k in rng.choice(np.flatnonzero(free), 20, replace=False):
        p1, p2 = pos.copy(), pos.copy()
        p1[k] += h
        p2[k] -= h
        g = (_energy(n, 1.0, 1.0, p1) - _energy(n, 1.0, 1.0, p2)) / (2 * h)
        assert abs(f[k] + g) <= 1e-6 * max(1.0, abs(g))
    assert np.all(f[~free] == 0)


def test_newton_third_law_without_clamps():
    n = O.Net.lattice(10, 0.7, 5, True).triples()
    pos = n.pos() + 0.03 * np.random.RandomState(1).randn(2 * n.n)
    f = n.forces(1.0, 1e-2, pos).reshape(-1, 2)
    assert np.abs(f.sum(axis=0)).max() < 1e-14


def test_single_bond_closed_form():
    n = O.Net.from_arrays([[0, 0], [1.3, 0.4]], [[0, 1]], [[0, 0]], 100.0, 100.0)
    L = math.hypot(1.3, 0.4)
    pos = np.array([0, 0, 1.3 * 1.5, 0.4 * 1.5])
    f = n.forces(2.0, 0.0, pos).reshape(-1, 2)
    d = 1.5 * L
    mag = 2.0 * (d - L)
    np.testing.assert_allclose(f[0], mag * np.array([1.3, 0.4]) / L, rtol=1e-14)
    np.testing.assert_allclose(f[1], -f[0], rtol=1e-14)
    es, eb = n.energies(2.0, 0.0, pos)
    assert abs(es - 0.5 * 2.0 * (d - L) ** 2) < 1e-15


def test_single_bent_fiber_closed_form():
    # straight 3-node fiber along x; end b rotated by phi about the hinge
    n = O.Net.from_arrays([[0, 0], [-1, 0], [1, 0]], [[0, 1], [0, 2]], [[0, 0], [0, 0]],
                          100.0, 100.0).triples()
    assert n.t == 1
    phi, r = 0.2, 1.7
    pos = np.array([0, 0, -1, 0, r * math.cos(phi), r * math.sin(phi)])
    kb = 0.3
    f = n.forces(0.0, kb, pos).reshape(-1, 2)
    # E = kb/2 phi^2 with phi = theta_b - theta_a - pi; each end feels
    # -dE/dr = -kb*phi * d(phi)/dr, of size kb*phi/|arm|, perpendicular to its
    # arm, both ends turning the fiber straight; the hinge balances them
    tang = np.array([-math.sin(phi), math.cos(phi)])
    np.testing.assert_allclose(f[2], -kb * phi / r * tang, rtol=1e-13)
    np.testing.assert_allclose(f[1], [0, -kb * phi / 1.0], rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(f[0], -f[1] - f[2], rtol=1e-13)
    es, eb = n.energies(0.0, kb, pos)
    assert abs(eb - 0.5 * kb * phi ** 2) < 1e-15


def test_energy_invariant_under_rigid_motion():
    n = O.Net.lattice(8, 0.75, 6, False).triples()
    pos = n.pos() + 0.05 * np.random.RandomState(2).randn(2 * n.n)
    e0 = _energy(n, 1.0, 1e-2, pos)
    moved = pos.reshape(-1, 2) + np.array([3.25, -1.5])
    assert abs(_energy(n, 1.0, 1e-2, moved.reshape(-1)) - e0) <= 1e-12 * e0


def test_fiber_kappa_constant_along_fibers():
    n = O.Net.lattice(12, 0.8, 2, True).triples().fiber_kappa(1e-5, 1e-2, 7)
    k = n.tri_kappa()
    lab = n.fiber_labels()
    for l in np.unique(lab):
        assert np.all(k[lab == l] == k[lab == l][0])
    assert k.min() >= 1e-5 and k.max() <= 1e-2
    assert len(np.unique(lab)) > 10


# ---- affine step and the full-lattice closed form --------------------------

def _full_lattice_energy_per_node(gamma):
    s = math.sqrt(3.0) / 2
    d60 = math.sqrt(1 + 2 * s * gamma * 0.5 + (s * gamma) ** 2)
    d120 = math.sqrt(1 - 2 * s * gamma * 0.5 + (s * gamma) ** 2)
    return 0.5 * ((d60 - 1) ** 2 + (d120 - 1) ** 2)


@pytest.mark.parametrize("gamma", [0.01, 0.2])
def test_full_lattice_affine_shear_is_equilibrium(gamma):
    n = O.Net.lattice(16, 1.0, 1, True).triples()
    pos = n.pos()
    n.shear(pos, gamma)
    f = n.forces(1.0, 1e-3, pos)
    assert np.abs(f).max() < 1e-12
    es, eb = n.energies(1.0, 1e-3, pos)
    assert abs(es / n.n - _full_lattice_energy_per_node(gamma)) <= 1e-12 * es / n.n
    assert eb < 1e-25
    r = n.fire(1.0, 1e-3, pos, fcut=1e-10)
    assert r["steps"] == 1 and r["rc"] == 0


def test_affine_matches_reference_expression():
    n = O.Net.lattice(6, 1.0, 1, False)
    pos = n.pos()
    miny, maxy = pos[1::2].min(), pos[1::2].max()
    yr = maxy - miny
    exp = pos.copy()
    exp[0::2] += (0.1 * yr) * (pos[1::2] - miny) / yr
    n.affine(pos, 0.1 * yr, 0.0)
    np.testing.assert_array_equal(pos, exp)


# ---- FIRE (run_fire_md :596-762) -------------------------------------------

def _chain_network():
    # clamped bottom (0,0) and top (0,2); free middle node at (0,1); rest lengths 1
    return O.Net.from_arrays([[0, 0], [0, 1], [0, 2]], [[0, 1], [1, 2]], [[0, 0], [0, 0]], 100.0)


def test_fire_first_steps_hand_computed():
    """1-D oscillator with k_eff = 2: the middle node starts at y = 1.1.
    Hand computation with dt = 0.1, m = 1 (drift, half kick, force, half kick):
      step 1: y = 1.1 + 0.005*(-0.2) = 1.099, v = -0.01 - 0.05*0.198 = -0.0199
      step 2: y = 1.099 - 0.1*0.0199 - 0.005*0.198 = 1.09602
    The mix leaves a velocity parallel to F unchanged, and P > 0 keeps dt."""
    n = _chain_network()
    pos = np.array([0, 0, 0, 1.1, 0, 2])
    r = n.fire(1.0, 0.0, pos, fp=(0.1, 0.99, 0.1, 1.1, 0.5), nmin=5, mass=1.0, fcut=0.0,
               max_steps=1)
    assert r["steps"] == 1
    assert abs(pos[3] - 1.099) < 1e-15
    assert abs(r["rms"] - 0.198 / math.sqrt(6)) < 1e-15
    pos = np.array([0, 0, 0, 1.1, 0, 2])
    n.fire(1.0, 0.0, pos, fp=(0.1, 0.99, 0.1, 1.1, 0.5), nmin=5, mass=1.0, fcut=0.0,
           max_steps=2)
    assert abs(pos[3] - 1.09602) < 1e-15
    assert pos[0] == 0 and pos[1] == 0 and pos[5] == 2  # clamped nodes never move


def test_fire_converges_to_independent_minimum():
    """FIRE's relaxed energy equals L-BFGS (scipy) on the same energy."""
    scipy_opt = pytest.importorskip("scipy.optimize")
    n = O.Net.lattice(8, 0.75, 3, True).triples()
    pos = n.pos()
    n.shear(pos, 0.02)
    x0 = pos.copy()
    r = n.fire(1.0, 1e-2, pos, fp=(0.1, 0.99, 0.1, 1.1, 0.5), fcut=1e-11)
    res = scipy_opt.minimize(lambda x: _energy(n, 1.0, 1e-2, x), x0,
                             jac=lambda x: -n.forces(1.0, 1e-2, x), method="L-BFGS-B",
                             options=dict(gtol=1e-13, ftol=1e-16, maxiter=100000))
    e_fire = r["es"] + r["eb"]
    assert abs(e_fire - res.fun) <= 1e-10 * e_fire


def test_fire_rms_vs_max_criterion():
    n = O.Net.lattice(8, 0.7, 8, True).triples()
    pos = n.pos()
    n.shear(pos, 0.02)
    r = n.fire(1.0, 1e-2, pos.copy(), fcut=1e-8, criterion=O.CRIT_RMS)
    assert r["rms"] < 1e-8
    m = n.fire(1.0, 1e-2, pos.copy(), fcut=1e-8, criterion=O.CRIT_MAX)
    assert m["fmax"] < 1e-8 and m["steps"] >= r["steps"]


def test_fire_reports_non_convergence_at_step_limit():
    n = O.Net.lattice(8, 0.7, 8, True).triples()
    pos = n.pos()
    n.shear(pos, 0.02)
    r = n.fire(1.0, 1e-2, pos, fcut=1e-12, max_steps=5)
    assert r["steps"] == 5 and r["fmax"] > 1e-12


# ---- CG (extension) ---------------------------------------------------------

@pytest.mark.parametrize("pr", [False, True])
def test_cg_agrees_with_fire(pr):
    n = O.Net.lattice(8, 0.75, 3, True).triples()
    pos = n.pos()
    n.shear(pos, 0.02)
    p1, p2 = pos.copy(), pos.copy()
    rf = n.fire(1.0, 1e-2, p1, fcut=1e-11)
    rc = n.cg(1.0, 1e-2, p2, fcut=1e-11, pr=pr)
    assert rc["rc"] == 0 and rc["fmax"] < 1e-11
    assert abs((rf["es"] + rf["eb"]) - (rc["es"] + rc["eb"])) <= 1e-10 * (rf["es"] + rf["eb"])


@pytest.mark.parametrize("pr", [False, True])
def test_cg_quadratic_chain_converges_in_n_steps(pr):
    """Collinear chain stretched along x between clamped ends: the energy is
    quadratic in the x displacements of the free nodes, so CG with an exact
    (secant) line search reaches the minimum in at most that many steps; on a
    quadratic successive forces are orthogonal, so Polak-Ribiere+ takes the
    Fletcher-Reeves directions and the same bound holds for it."""
    m = 6
    pos = [[i * 1.0, 0.0 if i in (0, m) else 1.0] for i in range(m + 1)]
    # y of the end nodes is 0 and 2 -> clamps; free nodes at y = 1 but then
    # bonds are not collinear, so lay the chain along y instead
    pos = [[0.0, float(i)] for i in range(m + 1)]
    bonds = [[i, i + 1] for i in range(m)]
    n = O.Net.from_arrays(pos, bonds, [[0, 0]] * m, 100.0)
    p = np.array(pos, float).reshape(-1).copy()
    p[1::2][1:m] += 0.1 * np.sin(np.arange(1, m))
    p[2 * m + 1] = m * 1.2   # stretched top end
    r = n.cg(1.0, 0.0, p, fcut=1e-12, criterion=O.CRIT_MAX, pr=pr)
    assert r["rc"] == 0 and r["steps"] <= m - 1 + 1
    np.testing.assert_allclose(p[1::2], np.arange(m + 1) * 1.2, atol=1e-11)


def _off_direction(s, d):
    """|component of s orthogonal to d| / |s|."""
    u = d / np.linalg.norm(d)
    return np.linalg.norm(s - (s @ u) * u) / np.linalg.norm(s)


@pytest.mark.parametrize("pr", [False, True])
def test_cg_second_direction_is_the_textbook_one(pr):
    """On a sheared, bent network (energy far from quadratic) the second CG
    step must lie along F1 + beta d0 with d0 = F0 and the textbook beta:
    Fletcher-Reeves |F1|^2/|F0|^2, Polak-Ribiere+ max(0, F1.(F1 - F0)/|F0|^2).
    The points come from the oracle's CG stopped after one and after two
    iterations and the forces from its force routine (pinned by finite
    differences); the line search leaves F1.F0 small but not zero, so the
    two betas give directions 6e-4 apart (relative) while the method's own
    is met to 1e-13, and the other method's direction (or a sign error in
    the numerator) is rejected."""
    n = O.Net.lattice(10, 0.8, 11, True).triples()
    x0 = n.pos()
    n.shear(x0, 0.08)
    ks, kb = 1.0, 5e-2
    f0 = n.forces(ks, kb, x0)
    x1 = x0.copy()
    assert n.cg(ks, kb, x1, fcut=0.0, max_iters=1, pr=pr)["steps"] == 1
    x2 = x0.copy()
    assert n.cg(ks, kb, x2, fcut=0.0, max_iters=2, pr=pr)["steps"] == 2
    f1 = n.forces(ks, kb, x1)
    # the first step is steepest descent
    assert _off_direction(x1 - x0, f0) < 1e-10
    b_fr = (f1 @ f1) / (f0 @ f0)
    b_pr = max(0.0, f1 @ (f1 - f0) / (f0 @ f0))
    b_sign = max(0.0, f1 @ (f1 + f0) / (f0 @ f0))
    assert b_pr > 0.0 and abs(b_fr - b_pr) > 1e-3 * b_fr  # the methods differ here
    step = x2 - x1
    want, other = (b_pr, b_fr) if pr else (b_fr, b_pr)
    assert _off_direction(step, f1 + want * f0) < 1e-9
    assert _off_direction(step, f1 + other * f0) > 1e-6
    assert _off_direction(step, f1 + b_sign * f0) > 1e-6


# ---- batch file and output names (:188-213, :297-383) ----------------------

def test_report_name_padding():
    assert O.report_name("b", 3, 7, "pos") == "b_007.pos"
    assert O.report_name("net_t", 2, 12, "pos") == "net_t_12.pos"
    assert O.dec_digits(0) == 1 and O.dec_digits(9) == 1 and O.dec_digits(100) == 3


def test_parse_batch_file(tmp_path):
    p = _write(tmp_path, "b.txt", "# comment\n\nnet tag\n1 0.001 1e-5 1e-2 7\n"
               "0.1 0.99 0.1 1.1 0.5\n5 1 1e-10 100000 max cg\n0 0 0\n0.01 0.03 0.01\n")
    rc, base, tag, b = O.parse_batch(p)
    assert rc == 0 and base == "net" and tag == "tag"
    assert (b.ks, b.kb, b.kappa_lo, b.kappa_hi, b.kappa_seed) == (1, 0.001, 1e-5, 1e-2, 7)
    assert list(b.fp) == [0.1, 0.99, 0.1, 1.1, 0.5]
    assert (b.nmin, b.mass, b.fcut, b.nmax, b.criterion, b.method) == (5, 1, 1e-10, 100000, 1, 1)
    assert (b.smin, b.smax, b.sinc) == (0.01, 0.03, 0.01)


@pytest.mark.parametrize("text", [
    "net\n1\n0.1 0.99 0.1 1.1 0.5\n5 1 1e-10 10\n0 0 0\n0 0 0\n",          # ks kb
    "net\n1 0\n0.1 0.99 0.1 1.1\n5 1 1e-10 10\n0 0 0\n0 0 0\n",            # 4 FIRE params
    "net\n1 0\n0.1 0.99 0.1 1.1 0.5\n5 1 1e-10\n0 0 0\n0 0 0\n",           # line 4
    "net\n1 0\n0.1 0.99 0.1 1.1 0.5\n5 1 1e-10 10\n0.2 0.1 0.1\n0 0 0\n",  # max < min
    "net\n1 0\n0.1 0.99 0.1 1.1 0.5\n5 1 1e-10 10\n0 0 0\n",               # no shear line
])
def test_parse_batch_file_rejections(tmp_path, text):
    assert O.parse_batch(_write(tmp_path, "b.txt", text))[0] != 0


def test_protocol_outputs_formats(tmp_path):
    n = O.Net.lattice(8, 0.8, 3, False)
    dat = str(tmp_path / "net1.dat")
    with open(dat, "w") as f:
        pos = n.pos().reshape(-1, 2)
        f.write(f"{n.n} {n.e} 8\n")
        for x, y in pos:
            f.write(f"{float(x)!r} {float(y)!r}\n")
        for (i, j), (w, _) in zip(n.bonds(), n.bond_wraps()):
            f.write(f"{i} {j} {w}\n")
    bf = _write(tmp_path, "b.txt", "net1 t\n1 0.01\n0.1 0.99 0.1 1.1 0.5\n5 1 1e-9 100000\n"
                "0 0 0\n0.01 0.02 0.01\n")
    rc, base, tag, b = O.parse_batch(bf)
    assert rc == 0
    base_out = str(tmp_path / "net1_t")
    assert O.run_protocol(dat, base_out + "_log.txt", base_out + "_edata.txt", base_out, b) == 0
    log = open(base_out + "_log.txt").read().splitlines()
    assert log[0].startswith("Convergence reached: Steps: ")
    assert log[-1].startswith("Elapsed time: ")
    ed = open(base_out + "_edata.txt").read().splitlines()
    assert len(ed) == 2
    cols = ed[0].split("\t")
    assert len(cols) == 5 and "e" in cols[2]
    assert float(cols[3]) >= 0 and abs(float(cols[4]) - float(cols[2]) - float(cols[3])) < 1e-9
    lines = open(base_out + "_1.pos").read().splitlines()
    assert len(lines) == n.n and "\t" in lines[0]


@pytest.mark.parametrize("le", [True, False])
def test_shear_virial_is_the_energy_derivative(le):
    """The shear virial equals dE/dgamma (central difference of the energy
    under the affine shear step, image shift included)."""
    o = O.Net.lattice(14, 0.72, 21, le).triples().fiber_kappa(1e-3, 1e-1, 5)
    pos = o.pos()
    rng = np.random.RandomState(3)
    pos += 0.05 * rng.standard_normal(pos.shape)
    o.shear(pos, 0.03)
    w = o.shear_virial(1.0, 1.0, pos)
    eps = 1e-5
    p = pos.copy()
    o.shear(p, eps)
    ep = sum(o.energies(1.0, 1.0, p))
    o.shear(p, -2 * eps)
    em = sum(o.energies(1.0, 1.0, p))
    o.shear(p, eps)
    fd = (ep - em) / (2 * eps)
    assert abs(w - fd) <= 1e-6 * abs(w), (w, fd)


def test_shear_virial_full_lattice_closed_form():
    """Full lattice (p = 1) under affine shear gamma: W = dE/dgamma of
    E = N ks/2 [(d60 - 1)^2 + (d120 - 1)^2], d^2 = 1 +- s gamma + s^2 gamma^2,
    s = sqrt(3)/2 (bending terms vanish: fibers stay straight)."""
    L, g = 16, 0.04
    o = O.Net.lattice(L, 1.0, 1, True).triples()
    pos = o.pos()
    o.shear(pos, g)
    s = np.sqrt(3.0) / 2
    d60 = np.sqrt(1 + s * g + (s * g) ** 2)
    d120 = np.sqrt(1 - s * g + (s * g) ** 2)
    dd60 = (s + 2 * s * s * g) / (2 * d60)
    dd120 = (-s + 2 * s * s * g) / (2 * d120)
    w_exact = L * L * ((d60 - 1) * dd60 + (d120 - 1) * dd120)
    w = o.shear_virial(1.0, 1e-3, pos)
    assert abs(w - w_exact) <= 1e-12 * abs(w_exact), (w, w_exact)
