"""Input discovery and output naming of the batch protocol: get_file_lists
(file_lists.cpp:80-137) and report_name (large_strain_relax.cu:203-213), seen
through nr_xml_job_plan, which lists the networks a job would run and their
output names without a GPU."""
import os

import pytest

import network_relaxation_amd as N

FIRE_PARS = "0.1 0.99 0.1 1.1 0.5 5 1.0 1e-10 1000000"


def _plan(path, base, tag="t"):
    job = os.path.join(path, "job.xml")
    with open(job, "w") as f:
        f.write(f'<rbatch><name input="{base}" tag="{tag}"/><mech_params stretch="1" bend="0"/>'
                f'<rscheme type="FIRE" pars="{FIRE_PARS}"/>'
                '<run><shear min="0" max="0.01" inc="0.01"/></run></rbatch>')
    lines = N.xml_job_plan(job).splitlines()
    return [ln.split() for ln in lines if ln.startswith("network ")]


def _touch(*paths):
    for p in paths:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p, "w") as f:
            f.write("0 0 1\n")


def test_directory_from_tokens(tmp_path, monkeypatch):
    """file_lists.cpp:93-103: '/'-split tokens, empty ones dropped; with two
    or more tokens all but the last rebuild the directory (leading '/' kept),
    with one the directory is "./"."""
    d = str(tmp_path)
    _touch(f"{d}/a/net1.dat", f"{d}/a/other.dat", f"{d}/x1.dat", f"{d}/sub_q.dat",
           f"{d}/sub/inner.dat")
    monkeypatch.chdir(d)
    # "a/net": directory a/, base net
    assert [r[1] for r in _plan(d, "a/net")] == ["a/net1.dat"]
    # absolute directory keeps its leading '/'
    assert [r[1] for r in _plan(d, f"{d}/a/net")] == [f"{d}/a/net1.dat"]
    # "/x" is one token: the reference searches ./ for x*, not /
    assert [r[1] for r in _plan(d, "/x")] == ["./x1.dat"]
    # "sub/" is one token: ./sub*.dat, not every .dat inside sub/
    assert [r[1] for r in _plan(d, "sub/")] == ["./sub_q.dat"]
    # doubled slashes are empty tokens
    assert [r[1] for r in _plan(d, "a//net")] == ["a/net1.dat"]
    # "*" matches every .dat of the directory
    assert [r[1] for r in _plan(d, "a/*")] == ["a/net1.dat", "a/other.dat"]


def test_prefix_names_stay_paired(tmp_path):
    """net1 and net10 both match base "net"; each input keeps its own output
    names (the reference sorts its four lists independently, and
    "net10.dat" < "net1_t_log.txt" pairs net1.dat with net10's outputs)."""
    d = str(tmp_path)
    _touch(f"{d}/net1.dat", f"{d}/net10.dat", f"{d}/net2.dat")
    rows = _plan(d, f"{d}/net")
    assert [r[1] for r in rows] == [f"{d}/net1.dat", f"{d}/net10.dat", f"{d}/net2.dat"]
    for r in rows:
        stem = os.path.basename(r[1])[:-4]
        assert r[5] == f"{d}/{stem}_t_log.txt"
        assert r[7] == f"{d}/{stem}_t_edata.txt"
        assert r[9] == f"{d}/{stem}_t"


def test_no_tag_names(tmp_path):
    """get_log_name / get_edat_name / get_disp_name (file_lists.cpp:17-50)
    without a tag: <m>_log.txt, <m>_edata.txt, <m>."""
    d = str(tmp_path)
    _touch(f"{d}/q.dat")
    job = os.path.join(d, "job.xml")
    with open(job, "w") as f:
        f.write(f'<rbatch><name input="{d}/q"/><mech_params stretch="1" bend="0"/>'
                f'<rscheme type="FIRE" pars="{FIRE_PARS}"/>'
                '<run><shear min="0" max="0.01" inc="0.01"/></run></rbatch>')
    row = [ln.split() for ln in N.xml_job_plan(job).splitlines() if ln.startswith("network")][0]
    assert row[4:] == ["log", f"{d}/q_log.txt", "edata", f"{d}/q_edata.txt", "disp", f"{d}/q"]


def test_only_the_extension_after_a_dot(tmp_path):
    """The pattern ^(<base>.*)\\.dat is searched, not matched whole: the
    stem is everything before the first ".dat" the search finds."""
    d = str(tmp_path)
    _touch(f"{d}/n.dat.bak", f"{d}/nn.txt")
    rows = _plan(d, f"{d}/n")
    assert [r[1] for r in rows] == [f"{d}/n.dat"]
    assert rows[0][5] == f"{d}/n_t_log.txt"
