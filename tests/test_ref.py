"""The reference's own host functions against the oracle and the product:
read_network, get_triples, get_dec_digits and report_name
(large_strain_relax.cu:42-213), compiled from the reference's source by
`make -C oracle ref` (oracle/_ref/libref_host.so; only where the reference
is present: the test skips elsewhere).  Points, bonds, wraps, rest lengths,
y extremes, triple order and wraps, accept/reject decisions and report names
must be identical bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

REF_SO = os.path.join(os.path.dirname(O.__file__), "_ref", "libref_host.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF_SO),
                                reason="oracle/_ref not built (needs the reference's source)")

_ref = None


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        _ref.ref_load.restype = C.c_void_p
        _ref.ref_load.argtypes = [C.c_char_p, C.c_int]
        _ref.ref_sizes.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 3 + [C.POINTER(C.c_double)]
        _ref.ref_arrays.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        _ref.ref_free.argtypes = [C.c_void_p]
        _ref.ref_report_name.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.c_char_p,
                                         C.c_int]
    return _ref


def ref_read(path, triples=True):
    h = ref().ref_load(os.fsencode(path), 1 if triples else 0)
    if not h:
        return None
    n, e, t = C.c_int(), C.c_int(), C.c_int()
    info = (C.c_double * 3)()
    ref().ref_sizes(h, C.byref(n), C.byref(e), C.byref(t), info)
    pos = np.zeros(2 * n.value)
    bonds = np.zeros(2 * max(e.value, 1), np.int32)
    rest = np.zeros(max(e.value, 1))
    wrap = np.zeros(max(e.value, 1), np.int8)
    tri = np.zeros(3 * max(t.value, 1), np.int32)
    tw = np.zeros(2 * max(t.value, 1), np.int8)
    ref().ref_arrays(h, pos.ctypes.data, bonds.ctypes.data, rest.ctypes.data, wrap.ctypes.data,
                     tri.ctypes.data, tw.ctypes.data)
    ref().ref_free(h)
    e, t = e.value, t.value
    return dict(pos=pos, bonds=bonds[:2 * e].reshape(-1, 2), rest=rest[:e], wrap=wrap[:e],
                tris=tri[:3 * t].reshape(-1, 3), twrap=tw[:2 * t].reshape(-1, 2),
                offset=info[0], miny=info[1], maxy=info[2])


def _compare(path):
    r = ref_read(path)
    o = O.Net.read(path).triples()
    p = N.Network.read(path).triples()
    pa, pi = p.arrays(), p.info()
    for name, pos in (("oracle", o.pos()), ("product", pa["pos"])):
        np.testing.assert_array_equal(pos, r["pos"], err_msg=name)
    np.testing.assert_array_equal(o.bonds(), r["bonds"])
    np.testing.assert_array_equal(pa["bonds"], r["bonds"])
    np.testing.assert_array_equal(o.bond_wraps()[:, 0], r["wrap"])
    np.testing.assert_array_equal(pa["bond_wraps"][:, 0], r["wrap"])
    # rest lengths bit for bit (sqrt(dx*dx + dy*dy), :97-99)
    assert o.rest().tobytes() == r["rest"].tobytes()
    assert pa["rest"].tobytes() == r["rest"].tobytes()
    assert (o.s.miny, o.s.maxy) == (r["miny"], r["maxy"])
    assert (pi.miny, pi.maxy, pi.width) == (r["miny"], r["maxy"], r["offset"])
    # triples: same sequence (hinges ascending, neighbour pairs in adjacency
    # order) and the x wraps of both arms
    np.testing.assert_array_equal(o.tris(), r["tris"])
    np.testing.assert_array_equal(pa["triples"], r["tris"])
    np.testing.assert_array_equal(o.tri_wraps()[:, [0, 2]], r["twrap"])
    np.testing.assert_array_equal(pa["triple_wraps"][:, [0, 2]], r["twrap"])
    return r


@pytest.mark.parametrize("L,p,seed", [(6, 1.0, 1), (12, 0.7, 2), (24, 0.62, 3), (33, 0.8, 4),
                                      (48, 0.55, 5)])
def test_lattices_match_the_reference(tmp_path, L, p, seed):
    path = str(tmp_path / "net.dat")
    N.Network.lattice(L, p, seed, False).write_dat(path)
    r = _compare(path)
    assert len(r["tris"]) > 0 or p < 0.6


def test_near_collinear_pairs_and_jittered_points(tmp_path):
    """The collinearity test |dot^2 - |a|^2 |b|^2| < 1e-10 (:176) near its
    threshold, arms of several lengths, plus random points (no triples)."""
    rng = np.random.default_rng(11)
    lines, pos, bonds = [], [], []
    for k, d in enumerate([0.0, 3e-6, 9.9e-6, 1.01e-5, 3e-5, 1e-3]):
        for scale in (0.5, 1.0, 2.0):
            base = len(pos)
            x0 = 10.0 * k + 3.0 * scale
            pos += [(x0, 5.0 * scale), (x0 + scale, 5.0 * scale + d), (x0 - 1.5 * scale, 5.0 * scale)]
            bonds += [(base, base + 1, 0), (base + 2, base, 0)]
    base = len(pos)
    pos += [(float(a), float(b)) for a, b in rng.uniform(0, 50, size=(30, 2))]
    bonds += [(base + i, base + (i * 7 + 3) % 30, int(rng.integers(-1, 2))) for i in range(30)]
    path = str(tmp_path / "c.dat")
    with open(path, "w") as f:
        f.write(f"{len(pos)} {len(bonds)} 60.0\n")
        for x, y in pos:
            f.write(f"{x!r} {y!r}\n")
        for b in bonds:
            f.write("%d %d %d\n" % b)
    r = _compare(path)
    assert 0 < len(r["tris"]) < 18


def test_negative_y_keeps_flt_min_maximum(tmp_path):
    """max_y starts at FLT_MIN, a positive number (:53-54): a network whose
    points all have y <= 0 reports that, not its largest y."""
    path = str(tmp_path / "neg.dat")
    with open(path, "w") as f:
        f.write("3 2 4\n0 -1\n1 -2\n2 -1.5\n0 1 0\n1 2 0\n")
    r = _compare(path)
    assert r["maxy"] > 0 and r["maxy"] < 1e-37


@pytest.mark.parametrize("text", [
    "2 1\n0 0\n1 1\n0 1 0\n", "0 1 5\n", "2 1 -1\n0 0\n1 1\n0 1 0\n", "3 1 5\n0 0\n1 1\n",
    "2 1 5\n0 0\n1 1", "2 1 5\n0 0\n1 1\n0 2 0\n", "2 2 5\n0 0\n1 1\n0 1 0\n",
    "2 1 5\njunk\n0 0\n\n1 1\nnot an edge\n0 1 0\n", "2 1 5\n0 0\n1 1\n0 1 0",
    "2 1 5\n0 0 7\n1 1 x\n0 1 0 9\n",
])
def test_accept_and_reject_like_the_reference(tmp_path, text):
    path = str(tmp_path / "x.dat")
    with open(path, "w") as f:
        f.write(text)
    r = ref_read(path)
    try:
        O.Net.read(path)
        o_ok = True
    except ValueError:
        o_ok = False
    try:
        N.Network.read(path)
        p_ok = True
    except N.NRError:
        p_ok = False
    assert o_ok == p_ok == (r is not None)
    if r is not None:
        _compare(path)


def test_report_names_and_digits():
    out = C.create_string_buffer(512)
    for v in (0, 1, 9, 10, 99, 100, 12345):
        assert ref().ref_dec_digits(v) == O.dec_digits(v)
    for base, digits, count, ext in (("b", 3, 7, "pos"), ("dir/x_t", 2, 11, "pos"),
                                     ("q", 1, 123, "disp"), ("", 4, 0, "pos")):
        ref().ref_report_name(base.encode(), digits, count, ext.encode(), out, 512)
        assert out.value.decode() == O.report_name(base, digits, count, ext)
