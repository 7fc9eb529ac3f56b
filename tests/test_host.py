"""CPU tests of the product library's host side (no GPU needed): the C ABI
exports every function include/nrelax.h declares, and the host topology the
library builds (lattice generator, get_triples, per-fiber kappa, clamp codes,
.dat reader and writer) is bit-identical to the FP64 oracle's.
"""
import os
import re

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "nrelax.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    names = _declared()
    assert len(names) >= 30
    lib = N.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("L,p,seed,le", [(8, 0.7, 3, True), (12, 0.6, 9, False),
                                         (20, 1.0, 1, True), (17, 0.55, 4, False)])
def test_host_topology_bit_identical_to_oracle(L, p, seed, le):
    n = N.Network.lattice(L, p, seed, le).triples().fiber_kappa(1e-5, 1e-2, seed + 1)
    o = O.Net.lattice(L, p, seed, le).triples().fiber_kappa(1e-5, 1e-2, seed + 1)
    a = n.arrays()
    np.testing.assert_array_equal(a["pos"], o.pos())
    np.testing.assert_array_equal(a["bonds"], o.bonds())
    np.testing.assert_array_equal(a["bond_wraps"], o.bond_wraps())
    np.testing.assert_array_equal(a["rest"], o.rest())
    np.testing.assert_array_equal(a["triples"], o.tris())
    np.testing.assert_array_equal(a["triple_wraps"], o.tri_wraps())
    np.testing.assert_array_equal(a["triple_kappa"], o.tri_kappa())
    np.testing.assert_array_equal(a["fcodes"], o.fcode())


def test_dat_write_read_round_trip(tmp_path):
    n = N.Network.lattice(10, 0.75, 5, False).triples()
    path = str(tmp_path / "net.dat")
    n.write_dat(path)
    back = N.Network.read(path).triples()
    ref = O.Net.read(path).triples()
    a, b = n.arrays(), back.arrays()
    for key in ("pos", "bonds", "bond_wraps", "rest", "triples", "triple_wraps", "fcodes"):
        np.testing.assert_array_equal(a[key], b[key])
    np.testing.assert_array_equal(b["rest"], ref.rest())
    np.testing.assert_array_equal(b["triples"], ref.tris())


def test_reader_errors_match_reference_codes(tmp_path):
    bad = tmp_path / "bad.dat"
    bad.write_text("3 2\n")  # header needs N E width (large_strain_relax.cu:60-65)
    with pytest.raises(N.NRError) as e:
        N.Network.read(str(bad))
    assert e.value.code == N.E_HEADER
    short = tmp_path / "short.dat"
    short.write_text("3 1 4.0\n0 0\n1 0\n")  # too few points (:81-85)
    with pytest.raises(N.NRError) as e:
        N.Network.read(str(short))
    assert e.value.code == N.E_POINTS
    idx = tmp_path / "idx.dat"
    idx.write_text("2 1 4.0\n0 0\n1 0\n0 5 0\n")  # bond index out of range (:100-104)
    with pytest.raises(N.NRError) as e:
        N.Network.read(str(idx))
    assert e.value.code == N.E_INDEX
    with pytest.raises(N.NRError) as e:
        N.Network.read(str(tmp_path / "absent.dat"))
    assert e.value.code == N.E_OPEN


def test_linear_response_argument_errors():
    """Rejected before any device work: a periodic box has no top row, and
    bending needs the triples."""
    le = N.Network.lattice(8, 0.8, 1, True).triples()
    with pytest.raises(N.NRError):
        N.linear_response(le, 1.0, 1e-3, 0.01, affine=False)
    bare = N.Network.lattice(8, 0.8, 1, False)
    with pytest.raises(N.NRError):
        N.linear_response(bare, 1.0, 1e-3, 0.01)
    # a batch is refused as a whole when any of its jobs is invalid
    good = N.Network.lattice(8, 0.8, 1, False)
    with pytest.raises(N.NRError):
        N.linear_response_batch([dict(net=good, ks=1.0, kb=0.0, shear=0.01),
                                 dict(net=bare, ks=1.0, kb=1e-3, shear=0.01)])
    with pytest.raises(N.NRError):
        N.linear_response_batch([dict(net=good, ks=1.0, kb=0.0, shear=0.01, tol=-1.0)])
    assert N.linear_response_batch([]) == []


def test_linear_cli_rejects_bad_batch_files(tmp_path):
    """nrelax_linear (the sp_relax drop-in) fails on an unreadable or invalid
    batch file before any device work, with the reference's messages."""
    import subprocess
    res = subprocess.run([N.LINEAR_CLI_PATH, "-b", str(tmp_path / "missing.txt")],
                         capture_output=True, text=True, timeout=60)
    assert res.returncode == 1 and "could not be read" in res.stderr
    bad = tmp_path / "bad.txt"
    bad.write_text("base\n1.0 -1 0\n0 0 0\n0 0.1 0.01\n")
    res = subprocess.run([N.LINEAR_CLI_PATH, "-b", str(bad)], capture_output=True, text=True,
                         timeout=60)
    assert res.returncode == 1 and "mechanical attributes" in res.stderr
    res = subprocess.run([N.LINEAR_CLI_PATH], capture_output=True, text=True, timeout=60)
    assert res.returncode == 255
