"""GPU: the nrelax CLI (lsrelax drop-in) against the oracle's protocol run on
the same batch file and network (reference mode: x-periodic, clamped rows)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import network_relaxation_amd as N
import oracle as O

pytestmark = pytest.mark.gpu


def _write_dat(net, path):
    a = net.arrays()
    i = net.info()
    with open(path, "w") as f:
        f.write(f"{i.n_nodes} {i.n_bonds} {i.width!r}\n")
        for x, y in a["pos"].reshape(-1, 2):
            f.write(f"{float(x)!r} {float(y)!r}\n")
        for (p, q), (w, _) in zip(a["bonds"], a["bond_wraps"]):
            f.write(f"{p} {q} {w}\n")


@pytest.mark.parametrize("absolute", [False, True])
def test_cli_protocol_matches_oracle(tmp_path, absolute):
    """Strain mode (displacements scaled by the y range, the default) and -a
    (absolute displacements, strain_mode false: large_strain_relax.cu:837)."""
    print(f"case cli protocol absolute={absolute}", flush=True)
    gpu_dir, cpu_dir = tmp_path / "gpu", tmp_path / "cpu"
    gpu_dir.mkdir()
    cpu_dir.mkdir()
    net = N.Network.lattice(16, 0.8, 9, False)
    ramps = "0.1 0.2 0.1\n0.1 0.3 0.1\n" if absolute else "0.01 0.02 0.01\n0.01 0.03 0.01\n"
    batch = "netA t\n1 0.01\n0.1 0.99 0.1 1.1 0.5\n5 1 1e-10 200000\n" + ramps
    for d in (gpu_dir, cpu_dir):
        _write_dat(net, str(d / "netA.dat"))
        (d / "batch.txt").write_text(batch.replace("netA", str(d / "netA")))
    args = [N.CLI_PATH, "-b", str(gpu_dir / "batch.txt")] + (["-a"] if absolute else [])
    res = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    rc, base, tag, b = O.parse_batch(str(cpu_dir / "batch.txt"))
    assert rc == 0
    stem = str(cpu_dir / "netA_t")
    assert O.run_protocol(str(cpu_dir / "netA.dat"), stem + "_log.txt", stem + "_edata.txt",
                          stem, b, strain_mode=not absolute) == 0
    g_log = (gpu_dir / "netA_t_log.txt").read_text().splitlines()
    c_log = (cpu_dir / "netA_t_log.txt").read_text().splitlines()
    assert len(g_log) == len(c_log) == 6
    for gl, cl in zip(g_log[:-1], c_log[:-1]):
        assert gl.split("Steps")[0] == cl.split("Steps")[0]   # same convergence verdict
    g_ed = (gpu_dir / "netA_t_edata.txt").read_text().splitlines()
    c_ed = (cpu_dir / "netA_t_edata.txt").read_text().splitlines()
    assert len(g_ed) == len(c_ed) == 5
    for gl, cl in zip(g_ed, c_ed):
        g, c = [float(x) for x in gl.split("\t")], [float(x) for x in cl.split("\t")]
        assert g[:2] == c[:2] and gl.split("\t")[:2] == cl.split("\t")[:2]
        for a, e in zip(g[2:], c[2:]):
            assert abs(a - e) <= 1e-9 * abs(e) + 1e-18
    for k in (1, 2, 3):
        gp = np.loadtxt(gpu_dir / f"netA_t_{k}.pos")
        cp = np.loadtxt(cpu_dir / f"netA_t_{k}.pos")
        assert gp.shape == cp.shape
        np.testing.assert_allclose(gp, cp, rtol=0, atol=1e-8)


def test_batched_directory_equals_each_network_alone(tmp_path):
    """relaxation_batch relaxes every <base>*.dat of the directory in one
    device batch (large_strain_relax.cu:1038-1043 loops over them one at a
    time).  With tiles_per_member fixed, each network's arithmetic does not
    depend on the others: four networks of different heights (so different
    displacement units in strain mode) relaxed together write, byte for byte,
    the files each writes when it is the only network, Elapsed-time lines
    aside.  A file that cannot be read is reported and skipped."""
    print("case batched directory vs alone", flush=True)
    specs = [("netA", 16, 0.8, 9), ("netB", 24, 0.75, 3), ("netC", 12, 0.85, 5),
             ("netD", 20, 0.7, 8)]
    batch = ("{base} t\n1 0.01\n0.1 0.99 0.1 1.1 0.5\n5 1 1e-10 200000\n"
             "0.01 0.02 0.01\n0.01 0.03 0.01\n")
    alld = tmp_path / "all"
    alld.mkdir()
    for name, L, p, seed in specs:
        net = N.Network.lattice(L, p, seed, False)
        _write_dat(net, str(alld / f"{name}.dat"))
        one = tmp_path / name
        one.mkdir()
        _write_dat(net, str(one / f"{name}.dat"))
        (one / "b.txt").write_text(batch.format(base=one / name))
    (alld / "netE.dat").write_text("3 1 5\n0 0\n")  # too few points: skipped
    (alld / "b.txt").write_text(batch.format(base=alld / "net"))
    assert N.relaxation_batch(str(alld / "b.txt"), tiles_per_member=2) == N.E_POINTS
    for name, *_ in specs:
        assert N.relaxation_batch(str(tmp_path / name / "b.txt"), tiles_per_member=2) == 0
        files = sorted(f.name for f in (tmp_path / name).iterdir()
                       if f.name.startswith(name + "_t"))
        assert len(files) == 2 + 3
        for f in files:
            a = (alld / f).read_text().splitlines()
            b = (tmp_path / name / f).read_text().splitlines()
            if f.endswith("_log.txt"):
                assert a[-1].startswith("Elapsed time:") and b[-1].startswith("Elapsed time:")
                a, b = a[:-1], b[:-1]
            assert a == b, f
    assert not (alld / "netE_t_log.txt").exists()


@pytest.mark.parametrize("itemize", [False, True])
def test_linear_cli_matches_oracle(tmp_path, itemize):
    """nrelax_linear (the sp_relax drop-in, strain mode) against the oracle's
    restatement of the same small-strain solve at every load step."""
    print(f"case linear cli itemize={itemize}", flush=True)
    net = N.Network.lattice(14, 0.75, 5, False)
    _write_dat(net, str(tmp_path / "lin.dat"))
    (tmp_path / "batch.txt").write_text(f"{tmp_path / 'lin'} q\n1.0 0.01 0.0\n0 0 0\n0 0.02 0.01\n")
    args = [N.LINEAR_CLI_PATH, "-b", str(tmp_path / "batch.txt"), "-s"] + (["-i"] if itemize else [])
    res = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    o = O.Net.read(str(tmp_path / "lin.dat")).triples()
    yext = o.s.maxy - o.s.miny
    lines = (tmp_path / "lin_q_edata.txt").read_text().splitlines()
    log = (tmp_path / "lin_q_log.txt").read_text().splitlines()
    assert len(lines) == 3 and len(log) == 4
    assert all(l.startswith("Residual RMS force: ") for l in log[:3])
    assert log[3].startswith("Elapsed time: ")
    for k, line in enumerate(lines):
        cols = line.split("\t")
        assert len(cols) == (4 if itemize else 3)
        comp, shear = float(cols[0]), float(cols[1])
        assert comp == 0.0
        ref = o.linear_response(1.0, 0.01, shear / yext)
        if itemize:
            assert float(cols[2]) == pytest.approx(ref["es"], rel=1e-9, abs=1e-20)
            assert float(cols[3]) == pytest.approx(ref["eb"], rel=1e-8, abs=1e-20)
        else:
            assert float(cols[2]) == pytest.approx(ref["es"] + ref["eb"], rel=1e-9, abs=1e-20)
        u = np.loadtxt(tmp_path / f"lin_q_{k + 1}.disp")
        assert u.shape == (o.n, 2)
        np.testing.assert_allclose(u.reshape(-1), ref["u"], rtol=0, atol=1e-8)
