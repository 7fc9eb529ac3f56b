"""The oracle's network + FEM-gel composite (oracle/mesh.c, restating
src/network_mesh_relax.cu), and the product's host side of it (reader,
synthetic composite, twin pairs).

Pinned by what the mathematics fixes, not by retyping the formulas:
  * the gel energy of a uniform deformation in closed form: a dilation by s
    costs bmod A (s^2 - 1 - d)^2 / 2 and no shear energy; a simple shear by
    gamma costs smod A gamma^2 / 2 (+ bmod A d^2 / 2); rotations cost nothing;
  * forces are minus the gradient of the energies (central differences),
    masked on clamped points, twins receiving the derivative with respect to
    moving both together (duplicate_forces :754-779);
  * an affinely sheared gel is an exact equilibrium (uniform stress; twins
    close the boundary), so FIRE stops after its first step;
  * the preferred dilation of a full lattice in closed form (every bond
    stretched from 1 to s: E = ks (s - 1)^2 / 2 per bond);
  * the reference's triple quirk (triples found with the dilated offset on
    undilated points) against a brute-force enumeration;
  * FIRE's end state is a local minimum of the oracle's energy.
"""
import math
import os

import numpy as np
import pytest

import oracle as O
import network_relaxation_amd as N

FP = (0.1, 0.99, 0.1, 1.1, 0.5)


def _lattice_file(tmp_path, L, p, seed, name="c"):
    path = os.path.join(str(tmp_path), f"{name}.nmc")
    N.Mesh.lattice(L, p, seed).write_nmc(path)
    return path


def _free_mask(m):
    fc = m.fcode()
    mask = np.zeros(2 * m.n, bool)
    mask[0::2] = (fc & 1) != 0
    mask[1::2] = (fc & 2) != 0
    return mask


def test_reader_matches_generator(tmp_path):
    """The oracle's reader gives the generator's points, facets, areas and
    twins; facets tile the strip (their areas sum to its area)."""
    L = 8
    m = N.Mesh.lattice(L, 0.7, 5)
    path = os.path.join(str(tmp_path), "g.nmc")
    m.write_nmc(path)
    a = m.arrays()
    o = O.Mesh.read(path)
    np.testing.assert_array_equal(o.pos(), a["pos"])
    np.testing.assert_array_equal(o.facets(), a["facets"])
    np.testing.assert_array_equal(o.areas(), a["areas"])
    np.testing.assert_array_equal(o.pairs(), a["twins"])
    assert o.areas().sum() == pytest.approx(L * (L - 1) * math.sqrt(3) / 2, rel=1e-14)
    # every left-edge point (x = 0) has a twin at x = L, same y
    pos = o.pos().reshape(-1, 2)
    for p_, q_ in o.pairs():
        assert pos[p_, 0] == 0.0 and pos[q_, 0] == L and pos[p_, 1] == pos[q_, 1]
    assert len(o.pairs()) == L


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_reader_hand_file_twins_and_errors(tmp_path):
    """A hand-written unit square: two left points twin with the two right
    ones (the coincident copy of a left point is not paired again); the
    reader's failures (:114-219) each give an error."""
    good = ("5 1 2 1.0\n0 0\n1 0\n0 1\n1 1\n0 0\n"  # point 4 coincides with point 0
            "0 1 0\n"
            "0 1 2\n1 3 2\n"
            "0.5\n0.5\n")
    p = os.path.join(str(tmp_path), "sq.nmc")
    _write(p, good)
    o = O.Mesh.read(p)
    assert o.n == 5 and o.e == 1 and o.nf == 2
    assert o.pairs().tolist() == [[0, 1], [2, 3]]
    assert list(o.fcode()) == [0, 0, 0, 0, 0]  # bottom and top rows clamped
    bad = {
        "header": "5 1 2\n",
        "points": "5 1 2 1.0\n0 0\n1 0\n",
        "facets_missing": "3 1 2 1.0\n0 0\n1 0\n0 1\n0 1 0\n0 1 2\n",
        "index": "3 1 1 1.0\n0 0\n1 0\n0 1\n0 1 0\n0 1 7\n0.5\n",
        "area": "3 1 1 1.0\n0 0\n1 0\n0 1\n0 1 0\n0 1 2\n-0.5\n",
        "areas_missing": "3 1 2 1.0\n0 0\n1 0\n0 1\n0 1 0\n0 1 2\n0 2 1\n0.5\n",
    }
    for k, text in bad.items():
        q = os.path.join(str(tmp_path), f"{k}.nmc")
        _write(q, text)
        with pytest.raises(ValueError):
            O.Mesh.read(q)


def test_gel_uniform_dilation_and_shear_closed_forms(tmp_path):
    m = O.Mesh.read(_lattice_file(tmp_path, 8, 0.7, 2))
    pos0 = m.pos()
    A = m.areas().sum()
    x0, y0 = m.s.minx, m.s.net.miny
    for s, d, bm in [(1.1, 0.0, 2.0), (0.95, 0.3, 1.5), (1.0, 0.2, 0.7)]:
        p = pos0.copy()
        p[0::2] = x0 + s * (p[0::2] - x0)
        p[1::2] = y0 + s * (p[1::2] - y0)
        assert m.gel_energy(0.8, bm, d, p) == pytest.approx(bm * A * (s * s - 1 - d) ** 2 / 2,
                                                           rel=1e-11, abs=1e-13)
    for g in [0.01, 0.2]:
        p = pos0.copy()
        p[0::2] += g * (p[1::2] - y0)
        assert m.gel_energy(0.8, 1.3, 0.1, p) == pytest.approx(
            0.8 * A * g * g / 2 + 1.3 * A * 0.01 / 2, rel=1e-11)


def test_gel_rotation_and_translation_invariance(tmp_path):
    m = O.Mesh.read(_lattice_file(tmp_path, 8, 0.7, 3))
    rng = np.random.default_rng(1)
    p = m.pos() + 0.05 * rng.standard_normal(2 * m.n)
    e0 = m.gel_energy(0.6, 1.2, 0.15, p)
    th = 0.7
    c, s = math.cos(th), math.sin(th)
    q = p.copy()
    q[0::2] = c * p[0::2] - s * p[1::2] + 3.0
    q[1::2] = s * p[0::2] + c * p[1::2] - 1.0
    assert m.gel_energy(0.6, 1.2, 0.15, q) == pytest.approx(e0, rel=1e-11)


@pytest.mark.parametrize("dil", [0.0, 0.1])
def test_forces_are_minus_energy_gradient(tmp_path, dil):
    """Central differences of E_stretch + E_bend + E_gel at a perturbed
    state: free non-twin DOFs get -dE/dx, twins get -dE/d(shift of both),
    clamped DOFs get 0."""
    m = O.Mesh.read(_lattice_file(tmp_path, 6, 0.8, 4))
    ks, kb, bmod, smod = 1.0, 0.02, 0.9, 0.4
    pos, pd = m.setup(ks, kb, bmod, dil)
    assert m.t > 0
    rng = np.random.default_rng(2)
    pos = pos + 0.03 * rng.standard_normal(pos.size)
    f = m.forces(ks, kb, smod, bmod, pd, pos)
    free = _free_mask(m)
    twin_of = {}
    for a, b in m.pairs():
        twin_of[a], twin_of[b] = b, a
    h = 1e-6
    for i in range(2 * m.n):
        if not free[i]:
            assert f[i] == 0.0
            continue
        node, c = divmod(i, 2)
        idx = [i] + ([2 * twin_of[node] + c] if node in twin_of else [])
        ep, em = pos.copy(), pos.copy()
        ep[idx] += h
        em[idx] -= h
        g = (sum(m.energies(ks, kb, smod, bmod, pd, ep)) -
             sum(m.energies(ks, kb, smod, bmod, pd, em))) / (2 * h)
        assert f[i] == pytest.approx(-g, abs=2e-7 * max(1.0, abs(g)))
    for a, b in m.pairs():  # twins receive one force
        assert f[2 * a] == f[2 * b] and f[2 * a + 1] == f[2 * b + 1]


def test_sheared_gel_is_an_equilibrium(tmp_path):
    """Gel only (ks = kb = 0, d = 0): the affine shear is the minimum; FIRE
    takes one step and E = smod A gamma^2 / 2."""
    m = O.Mesh.read(_lattice_file(tmp_path, 10, 0.6, 6))
    pos, pd = m.setup(0.0, 0.0, 1.0, 0.0)
    assert pd == 0.0
    yr = m.s.net.maxy - m.s.net.miny
    g = 0.05
    m.affine(pos, g * yr, 0.0, yr)
    f = m.forces(0.0, 0.0, 0.7, 1.0, pd, pos)
    assert np.max(np.abs(f)) < 1e-13
    r = m.fire(0.0, 0.0, 0.7, 1.0, pd, pos, fcut=1e-10)
    assert r["rc"] == 0 and r["steps"] == 1
    assert r["eg"] == pytest.approx(0.7 * m.areas().sum() * g * g / 2, rel=1e-10)
    assert r["es"] == 0.0 and r["eb"] == 0.0


def test_preferred_dilation_closed_form(tmp_path):
    """Full lattice (p = 1): the dilation by s = sqrt(1 + d) stretches every
    bond from 1 to s, so net_bmod = ks (s - 1)^2 E / (area d^2) and the
    preferred dilation is d (bmod + net_bmod) / bmod; d = 0 gives 0."""
    L, ks, bmod, d = 8, 1.5, 2.0, 0.12
    m = O.Mesh.read(_lattice_file(tmp_path, L, 1.0, 1))
    s = math.sqrt(1 + d)
    area = L * (L - 1) * math.sqrt(3) / 2
    net_bmod = ks * (s - 1) ** 2 * m.e / (area * d * d)
    _, pd = m.setup(ks, 0.0, bmod, d)
    assert pd == pytest.approx(d * (bmod + net_bmod) / bmod, rel=1e-12)
    m2 = O.Mesh.read(_lattice_file(tmp_path, L, 1.0, 1, "c2"))
    assert m2.setup(ks, 0.0, bmod, 0.0)[1] == 0.0


def _brute_triples(pos, bonds, wraps, width):
    """Antiparallel neighbour pairs at each hinge, |dot^2 - |a|^2|b|^2| < 1e-10."""
    nbr = {}
    for (i, j), w in zip(bonds, wraps):
        nbr.setdefault(i, []).append((j, w))
        nbr.setdefault(j, []).append((i, -w))
    out = set()
    for h, lst in nbr.items():
        for a in range(len(lst)):
            for b in range(a + 1, len(lst)):
                (ja, wa), (jb, wb) = lst[a], lst[b]
                ax = pos[2 * ja] + width * wa - pos[2 * h]
                ay = pos[2 * ja + 1] - pos[2 * h + 1]
                bx = pos[2 * jb] + width * wb - pos[2 * h]
                by = pos[2 * jb + 1] - pos[2 * h + 1]
                dot = ax * bx + ay * by
                if dot < 0 and abs(dot * dot - (ax * ax + ay * ay) * (bx * bx + by * by)) < 1e-10:
                    out.add((h, ja, jb))
    return out


def test_triples_use_the_dilated_offset(tmp_path):
    """With dilation > 0 the reference finds triples with the grown offset on
    the undilated points: inclined fibers crossing the x boundary lose theirs
    (horizontal ones keep them)."""
    L = 8
    path = _lattice_file(tmp_path, L, 1.0, 1)
    m0 = O.Mesh.read(path)
    m0.setup(1.0, 1.0, 1.0, 0.0)
    m1 = O.Mesh.read(path)
    m1.setup(1.0, 1.0, 1.0, 0.2)
    pos, bonds = m0.pos(), m0.bonds()
    wraps = np.ctypeslib.as_array(m0.s.net.bwrap, shape=(2 * m0.e,)).reshape(-1, 2)[:, 0]
    t0 = {tuple(t) for t in m0.tris()}
    t1 = {tuple(t) for t in m1.tris()}
    assert t0 == _brute_triples(pos, bonds, wraps, float(L))
    assert t1 == _brute_triples(pos, bonds, wraps, L * math.sqrt(1.2))
    lost = t0 - t1
    assert len(lost) > 0 and t1 < t0
    for h, a, b in lost:  # each lost triple is inclined and crosses the boundary
        assert pos[2 * a + 1] != pos[2 * h + 1]
        assert abs(pos[2 * a] - pos[2 * b]) > 2.0


def test_fire_ends_at_a_local_minimum(tmp_path):
    m = O.Mesh.read(_lattice_file(tmp_path, 8, 0.65, 7))
    ks, kb, bmod, smod = 1.0, 1e-2, 0.5, 0.3
    pos, pd = m.setup(ks, kb, bmod, 0.05)
    yr = (m.s.net.maxy - m.s.net.miny) * math.sqrt(1.05)
    m.affine(pos, 0.03 * yr, 0.0, yr)
    e_start = sum(m.energies(ks, kb, smod, bmod, pd, pos))
    r = m.fire(ks, kb, smod, bmod, pd, pos, fcut=1e-11)
    assert r["rc"] == 0 and r["rms"] < 1e-11
    e = r["es"] + r["eb"] + r["eg"]
    assert e < e_start
    free = _free_mask(m)
    twin_of = {}
    for a, b in m.pairs():
        twin_of[a], twin_of[b] = b, a
    rng = np.random.default_rng(3)
    for _ in range(5):  # admissible perturbations: clamped points fixed, twins together
        dp = 1e-4 * rng.standard_normal(pos.size) * free
        for a, b in m.pairs():
            dp[2 * b:2 * b + 2] = dp[2 * a:2 * a + 2]
        assert sum(m.energies(ks, kb, smod, bmod, pd, pos + dp)) > e


def test_mesh_protocol_outputs(tmp_path):
    """or_run_mesh_protocol writes the reference's files: compression lines
    "%12.0lf\\t%12.10lf" + 4 energies, shear lines with the displacement
    (strain x height) and the last compression, one .pos per shear step."""
    d = str(tmp_path)
    path = _lattice_file(tmp_path, 6, 0.7, 8, "cmp")
    batch = os.path.join(d, "b.txt")
    _write(batch, f"{d}/cmp t\n1.0 0.01 0.8 0.4 0.05\n0.1 0.99 0.1 1.1 0.5\n"
                  "5 1.0 1e-9 200000\n0 0.02 0.01\n0.01 0.02 0.01\n")
    rc, base, tag, b = O.parse_mesh_batch(batch)
    assert rc == 0 and (b.ks, b.kb, b.bmod, b.smod, b.dilation) == (1.0, 0.01, 0.8, 0.4, 0.05)
    rc, pd = O.run_mesh_protocol(path, f"{d}/cmp_t_log.txt", f"{d}/cmp_t_edata.txt",
                                 f"{d}/cmp_t", b, True)
    assert rc == 0 and pd > 0.05
    lines = open(f"{d}/cmp_t_edata.txt").read().splitlines()
    assert len(lines) == 3 + 2
    m = O.Mesh.read(path)
    yr = (m.s.net.maxy - m.s.net.miny) * math.sqrt(1.05)
    comp = [float(x.split("\t")[1]) for x in lines[:3]]
    assert comp == pytest.approx([0.0, 0.01 * yr, 0.02 * yr], abs=1e-9)
    assert lines[0].split("\t")[0].strip() == "0"
    sh = [ln.split("\t") for ln in lines[3:]]
    assert [float(x[0]) for x in sh] == pytest.approx([0.01 * yr, 0.02 * yr], abs=1e-7)
    assert all(float(x[1]) == pytest.approx(0.02 * yr, abs=1e-7) for x in sh)
    for ln in lines:
        f = [float(x) for x in ln.split("\t")]
        assert len(f) == 6 and f[5] == pytest.approx(f[2] + f[3] + f[4], rel=1e-8)
    log = open(f"{d}/cmp_t_log.txt").read().splitlines()
    assert all(x.startswith("Convergence reached: Steps: ") for x in log[:5])
    assert log[5].startswith("Elapsed time: ") and log[5].endswith(" seconds")
    for k in (1, 2):
        assert os.path.exists(f"{d}/cmp_t_{k}.pos")
    assert not os.path.exists(f"{d}/cmp_t_3.pos")


def test_product_reader_agrees_with_oracle(tmp_path):
    """The product's .nmc reader (nr_mesh_read) gives the oracle's points,
    facets, areas and twins on the hand-written square, and rejects each bad
    file the oracle rejects, with the reader's codes."""
    good = ("5 1 2 1.0\n0 0\n1 0\n0 1\n1 1\n0 0\n0 1 0\n0 1 2\n1 3 2\n0.5\n0.5\n")
    p = os.path.join(str(tmp_path), "sq.nmc")
    _write(p, good)
    a = N.Mesh.read(p).arrays()
    o = O.Mesh.read(p)
    np.testing.assert_array_equal(a["pos"], o.pos())
    np.testing.assert_array_equal(a["facets"], o.facets())
    np.testing.assert_array_equal(a["areas"], o.areas())
    np.testing.assert_array_equal(a["twins"], o.pairs())
    bad = {
        "header": ("5 1 2\n", N.E_HEADER),
        "points": ("5 1 2 1.0\n0 0\n1 0\n", N.E_POINTS),
        "edges": ("3 2 1 1.0\n0 0\n1 0\n0 1\n0 1 0\n", N.E_EDGES),
        "facets_missing": ("3 1 2 1.0\n0 0\n1 0\n0 1\n0 1 0\n0 1 2\n", N.E_MESH),
        "index": ("3 1 1 1.0\n0 0\n1 0\n0 1\n0 1 0\n0 1 7\n0.5\n", N.E_INDEX),
        "area": ("3 1 1 1.0\n0 0\n1 0\n0 1\n0 1 0\n0 1 2\n-0.5\n", N.E_MESH),
        "areas_missing": ("3 1 2 1.0\n0 0\n1 0\n0 1\n0 1 0\n0 1 2\n0 2 1\n0.5\n", N.E_MESH),
    }
    for k, (text, code) in bad.items():
        q = os.path.join(str(tmp_path), f"{k}.nmc")
        _write(q, text)
        with pytest.raises(N.NRError) as e:
            N.Mesh.read(q)
        assert e.value.code == code, k
        with pytest.raises(ValueError):
            O.Mesh.read(q)


def test_composite_argument_errors(tmp_path):
    """Set-up rejects negative moduli or dilation before any device work."""
    m = N.Mesh.lattice(6, 0.7, 1)
    for args in [(-1.0, 0.0, 1.0, 0.0, 0.0), (1.0, -1.0, 1.0, 0.0, 0.0),
                 (1.0, 0.0, -1.0, 0.0, 0.0), (1.0, 0.0, 1.0, 0.0, -0.1),
                 (1.0, 0.0, 1.0, -0.5, 0.0)]:
        with pytest.raises(N.NRError) as e:
            N.Composite(m, *args)
        assert e.value.code == N.E_ARG
    with pytest.raises(N.NRError):
        N.Mesh.lattice(2, 0.5, 1)
