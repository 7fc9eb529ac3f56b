"""The XML job format that src/large_strain_types.hpp declares (elements
:36-39, children :42-44, attributes :48-50, the description :226-250) and
never parses: nr_xml_job_plan's reading of job files (no GPU work), and on
the GPU that a job equals the batch protocol it restates."""
import os
import subprocess

import numpy as np
import pytest

import network_relaxation_amd as N

FIRE_PARS = "0.1 0.99 0.1 1.1 0.5 5 1.0 1e-10 1000000"


def _job(tmp_path, body, name="job.xml"):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as f:
        f.write(body)
    return p


def _base(tmp_path, nets=1):
    d = str(tmp_path)
    for k in range(nets):
        N.Network.lattice(12, 0.7, 3 + k, False).write_dat(os.path.join(d, f"net{k}.dat"))
    return os.path.join(d, "net")


def test_plan_of_a_two_run_job(tmp_path):
    base = _base(tmp_path, nets=2)
    p = _job(tmp_path, f"""<?xml version="1.0"?>
<!-- a job with two runs -->
<rbatch>
  <name input="{base}" tag="t"/>
  <mech_params stretch="1.0" bend='2e-3'/>
  <rscheme type="FIRE" pars="{FIRE_PARS}"/>
  <run>
    <axial min="0" max="0.02" inc="0.01"/>
    <shear min="0" max="0" inc="0"/>
    <report disp="false"/>
  </run>
  <run><shear min="0.01" max="0.03" inc="0.01"/></run>
</rbatch>
""")
    plan = N.xml_job_plan(p).splitlines()
    assert plan[0] == f"input {base} tag t"
    assert plan[1] == "stretch 1 bend 0.002"
    assert plan[2] == "method FIRE pars 0.10000000000000001 0.98999999999999999 " \
                      "0.10000000000000001 1.1000000000000001 0.5 5 1 1e-10 1000000"
    assert plan[3] == "run 1 axial 0 0.02 0.01 shear 0 0 0 report edata log"
    assert plan[4] == "run 2 axial 0 0 0 shear 0.01 0.029999999999999999 0.01 report edata log disp"
    d = str(tmp_path)
    assert plan[5] == (f"network {d}/net0.dat run 1 log {d}/net0_t_run1_log.txt edata "
                       f"{d}/net0_t_run1_edata.txt disp {d}/net0_t_run1")
    assert plan[8] == (f"network {d}/net1.dat run 2 log {d}/net1_t_run2_log.txt edata "
                       f"{d}/net1_t_run2_edata.txt disp {d}/net1_t_run2")
    assert len(plan) == 9


def test_one_run_keeps_the_batch_names_and_conjgrad_pars(tmp_path):
    base = _base(tmp_path)
    p = _job(tmp_path, f"""<rbatch><name input="{base}"/><mech_params stretch="1" bend="0"/>
<rscheme type="ConjGrad" pars="1e-9 5000"/><run><shear min="0" max="0.01" inc="0.01"/></run>
</rbatch>""")
    plan = N.xml_job_plan(p).splitlines()
    d = str(tmp_path)
    assert plan[2].startswith("method ConjGrad pars")
    assert plan[2].endswith(" 1.0000000000000001e-09 5000")
    assert plan[4] == (f"network {d}/net0.dat run 1 log {d}/net0_log.txt edata "
                       f"{d}/net0_edata.txt disp {d}/net0")


@pytest.mark.parametrize("body,where", [
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/></rbatch>", "needs <name>"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/>"
     "<rscheme type='FIRE' pars='" + FIRE_PARS + "'/></rbatch>", "at least one <run>"),
    ("<rbatch><name input='x' colour='red'/></rbatch>", "no attribute colour"),
    ("<rbatch><axial min='0' max='1' inc='1'/></rbatch>", "cannot appear in <rbatch>"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='FIRE' "
     "pars='" + FIRE_PARS + "'/><run><shear min='0' max='1' inc='1'/><shear min='0' max='1' "
     "inc='1'/></run></rbatch>", "given twice"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='Newton' "
     "pars='1'/><run/></rbatch>", "FIRE or ConjGrad"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='FIRE' "
     "pars='0.1 0.99'/><run/></rbatch>", "need 9 numbers"),
    ("<rbatch><name input='x'/><mech_params stretch='one' bend='0'/><rscheme type='FIRE' "
     "pars='" + FIRE_PARS + "'/><run/></rbatch>", "not a finite number"),
    ("<rbatch><name input='x'/><mech_params stretch='nan' bend='0'/><rscheme type='FIRE' "
     "pars='" + FIRE_PARS + "'/><run/></rbatch>", "not a finite number"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='FIRE' "
     "pars='0.1 0.99 inf 1.1 0.5 5 1.0 1e-10 1000'/><run/></rbatch>", "not a finite number"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='FIRE' "
     "pars='0.1 0.99 0.1 1.1 0.5 1e12 1.0 1e-10 1000'/><run/></rbatch>", "must lie in"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='ConjGrad' "
     "pars='1e-10 -5'/><run/></rbatch>", "must lie in"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='FIRE' "
     "pars='" + FIRE_PARS + "'/><run><shear min='0' max='1e300' inc='1e-300'/></run></rbatch>",
     "invalid <shear> range"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='FIRE' "
     "pars='" + FIRE_PARS + "'/><run><axial min='0' max='inf' inc='0.1'/></run></rbatch>",
     "not a finite number"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='FIRE' "
     "pars='" + FIRE_PARS + "'/><run><shear min='0.5' max='0.1' inc='0.1'/></run></rbatch>",
     "invalid <shear> range"),
    ("<rbatch><name input='x'/><mech_params stretch='1' bend='0'/><rscheme type='FIRE' "
     "pars='" + FIRE_PARS + "'/><run><report log='maybe'/></run></rbatch>", "not true/false"),
    ("<rbatch><name input='x'>text</name></rbatch>", "text inside <name>"),
    ("<rbatch><name input='x'></rbatch>", "does not close <name>"),
    ("<job/>", "unknown element <job>"),
    ("<run/>", "root element must be <rbatch>"),
])
def test_format_errors(tmp_path, body, where):
    p = _job(tmp_path, body)
    with pytest.raises(N.NRError) as e:
        N.xml_job_plan(p)
    assert e.value.code == N.E_BATCH
    assert where in str(e.value)


def test_missing_file(tmp_path):
    with pytest.raises(N.NRError) as e:
        N.xml_job_plan(os.path.join(str(tmp_path), "absent.xml"))
    assert e.value.code == N.E_OPEN


@pytest.mark.gpu
def test_job_runs_continue_the_batch_protocol(tmp_path):
    """Two XML runs (shear 0.01..0.03, then 0.04..0.05, inc 0.01) on one
    network give, line for line, the edata of the 6-line batch with the whole
    ramp 0.01..0.05 (state carried between runs; same device path, bit for
    bit); a run with disp="false" writes no .pos files; `nrelax -j` runs it."""
    print("case xml job vs batch", flush=True)
    d = str(tmp_path)
    os.makedirs(f"{d}/xml")
    os.makedirs(f"{d}/bat")
    for sub in ("xml", "bat"):
        N.Network.lattice(12, 0.7, 3, False).triples().write_dat(f"{d}/{sub}/net0.dat")
    with open(f"{d}/bat/b.txt", "w") as f:
        f.write(f"{d}/bat/net t\n1.0 2e-3\n0.1 0.99 0.1 1.1 0.5\n5 1.0 1e-10 1000000\n"
                "0 0 0\n0.01 0.05 0.01\n")
    assert N.relaxation_batch(f"{d}/bat/b.txt") == 0
    job = _job(f"{d}/xml", f"""<rbatch><name input="{d}/xml/net" tag="t"/>
<mech_params stretch="1.0" bend="2e-3"/><rscheme type="FIRE" pars="{FIRE_PARS}"/>
<run><shear min="0.01" max="0.03" inc="0.01"/><report disp="false"/></run>
<run><shear min="0.04" max="0.05" inc="0.01"/></run></rbatch>""")
    out = subprocess.run([N.CLI_PATH, "-j", job], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    b = open(f"{d}/bat/net0_t_edata.txt").read().splitlines()
    r1 = open(f"{d}/xml/net0_t_run1_edata.txt").read().splitlines()
    r2 = open(f"{d}/xml/net0_t_run2_edata.txt").read().splitlines()
    assert b == r1 + r2
    assert not os.path.exists(f"{d}/xml/net0_t_run1_1.pos")
    for k in (1, 2):
        np.testing.assert_array_equal(np.loadtxt(f"{d}/xml/net0_t_run2_{k}.pos"),
                                      np.loadtxt(f"{d}/bat/net0_t_{k + 3}.pos"))


def test_mutated_job_files_fail_cleanly(tmp_path):
    """Random byte edits, truncations and insertions of a valid job file:
    the reader returns success or NR_E_BATCH, never anything else (and never
    crashes the process)."""
    base = _base(tmp_path)
    good = (f'<?xml version="1.0"?><rbatch><name input="{base}" tag="t"/>'
            f'<mech_params stretch="1.0" bend="2e-3"/>'
            f'<rscheme type="FIRE" pars="{FIRE_PARS}"/>'
            '<run><axial min="0" max="0.02" inc="0.01"/><shear min="0" max="0.1" inc="0.05"/>'
            '<report edata="true" log="false" disp="1"/></run></rbatch>')
    assert N.xml_job_plan(_job(tmp_path, good, "good.xml"))
    rng = np.random.default_rng(7)
    alphabet = b'<>/="\' \n!?-abcdefghijklmnopqrstuvwxyz0123456789.&;'
    for trial in range(400):
        b = bytearray(good.encode())
        for _ in range(int(rng.integers(1, 4))):
            op = int(rng.integers(0, 3))
            k = int(rng.integers(0, len(b)))
            if op == 0:
                b[k] = alphabet[int(rng.integers(0, len(alphabet)))]
            elif op == 1:
                del b[k:k + int(rng.integers(1, 8))]
            else:
                b[k:k] = bytes(alphabet[int(rng.integers(0, len(alphabet)))]
                               for _ in range(int(rng.integers(1, 6))))
        if trial % 50 == 0:
            b = b[:int(rng.integers(0, len(b)))]
        p = os.path.join(str(tmp_path), "m.xml")
        with open(p, "wb") as f:
            f.write(bytes(b))
        try:
            N.xml_job_plan(p)
        except N.NRError as e:
            assert e.code == N.E_BATCH, (e.code, bytes(b))


@pytest.mark.gpu
def test_conjgrad_job_equals_the_cg_batch(tmp_path):
    """rscheme type="ConjGrad" pars="fcut max_iters" runs the batch file's
    nonlinear CG (line 4 "... cg"): the same energy file line for line."""
    print("case xml ConjGrad job vs cg batch", flush=True)
    d = str(tmp_path)
    os.makedirs(f"{d}/xml")
    os.makedirs(f"{d}/bat")
    for sub in ("xml", "bat"):
        N.Network.lattice(10, 0.7, 4, False).triples().write_dat(f"{d}/{sub}/net0.dat")
    with open(f"{d}/bat/b.txt", "w") as f:
        f.write(f"{d}/bat/net t\n1.0 2e-3\n0.1 0.99 0.1 1.1 0.5\n5 1.0 1e-10 100000 rms cg\n"
                "0 0 0\n0.01 0.02 0.01\n")
    assert N.relaxation_batch(f"{d}/bat/b.txt") == 0
    job = _job(f"{d}/xml", f"""<rbatch><name input="{d}/xml/net" tag="t"/>
<mech_params stretch="1.0" bend="2e-3"/><rscheme type="ConjGrad" pars="1e-10 100000"/>
<run><shear min="0.01" max="0.02" inc="0.01"/></run></rbatch>""")
    assert N.xml_batch(job) == 0
    assert (open(f"{d}/bat/net0_t_edata.txt").read() ==
            open(f"{d}/xml/net0_t_edata.txt").read())
