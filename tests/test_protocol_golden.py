"""Golden outputs of the relax protocol (do_relaxation_run,
large_strain_relax.cu:792-1010), derived here without the oracle's writer.

The network is a full triangular lattice (every bond present; reference box:
x-periodic, clamped bottom and top rows) written by this file.  Any affine
deformation of a full lattice is an equilibrium (each free node is an
inversion centre of its six neighbours, and lattice lines stay straight), so
FIRE stops after its first step and every increment's energy is known in
closed form: with the bottom row fixed, c compression steps of strain c_inc
scale heights by s = (1 - c_inc)^c and m shear steps of strain s_inc give
x += m s_inc (y - y_min); a bond (dx, dy) becomes (dx + g dy', dy') with
dy' = s dy, g = m s_inc, so horizontal bonds keep length 1 and the 60 and
120 degree bonds (L (L - 1) of each) have
    d^2 = (+-1/2 + g s sqrt(3)/2)^2 + 3/4 s^2,
    E_stretch = ks/2 L (L - 1) [(d_60 - 1)^2 + (d_120 - 1)^2],  E_bend = 0
(kb = 0: no triples).  Expected file lines are formatted with the
reference's own printf strings: the log line (:948), the compression edata
line (:953), the shear edata line (:993), the .pos line (:223) and the
elapsed-time line (:1004); names from report_name (:203-213) with the padding
of get_dec_digits(1 + (int)((max - min) / inc)) (:838).  Both ramps start at
0, so each edata label lags the applied strain by one increment (the loop
applies `inc` before relaxing, :925-928, :967-970), and the shear lines carry
max_comp = the last compression label (:957).

The same expectations are checked for the oracle's protocol (CPU) and for the
nrelax CLI (GPU)."""
import math
import os
import re
import subprocess

import pytest

import network_relaxation_amd as N
import oracle as O

L = 8
KS = 1.0
CRAMP = (0.0, 0.02, 0.01)
SRAMP = (0.0, 0.1, 0.01)
BATCH = "{base} g\n%r 0\n0.1 0.99 0.1 1.1 0.5\n5 1.0 1e-10 100000\n%r %r %r\n%r %r %r\n" % (
    (KS,) + CRAMP + SRAMP)

LOG_FMT = "Steps: %d, RMS Force: %12.8le"          # :948
COMP_FMT = "%12.0lf\t%12.10lf\t%12.8le\t%12.8le\t%12.8le"     # :953
SHEAR_FMT = "%12.10lf\t%12.10lf\t%12.10le\t%12.10le\t%12.10le"  # :993
POS_FMT = "%2.10le\t%2.10le"                         # :223


def write_full_lattice(path):
    """L x L triangular lattice, spacing 1, rows sqrt(3)/2 apart, odd rows
    shifted by 1/2; bonds to the right neighbour (x-periodic) and to the two
    neighbours in the row above (none from the top row)."""
    h = math.sqrt(3.0) / 2.0
    pos, bonds = [], []
    for r in range(L):
        for c in range(L):
            pos.append((c + 0.5 * (r % 2), r * h))
    for r in range(L):
        for c in range(L):
            i = r * L + c
            bonds.append((i, r * L + (c + 1) % L, 1 if c == L - 1 else 0))
            if r == L - 1:
                continue
            if r % 2 == 0:
                bonds.append((i, (r + 1) * L + c, 0))                              # 60 deg
                bonds.append((i, (r + 1) * L + (c - 1) % L, -1 if c == 0 else 0))  # 120 deg
            else:
                bonds.append((i, (r + 1) * L + c, 0))                              # 120 deg
                bonds.append((i, (r + 1) * L + (c + 1) % L, 1 if c == L - 1 else 0))
    with open(path, "w") as f:
        f.write(f"{L * L} {len(bonds)} {float(L)!r}\n")
        for x, y in pos:
            f.write(f"{x!r} {y!r}\n")
        for i, j, w in bonds:
            f.write(f"{i} {j} {w}\n")
    return pos


def stretch_energy(s, g):
    h = math.sqrt(3.0) / 2.0
    e = 0.0
    for sx in (0.5, -0.5):
        d = math.hypot(sx + g * s * h, s * h)
        e += (d - 1.0) ** 2
    return 0.5 * KS * L * (L - 1) * e


def ramp_labels(lo, hi, inc, rng):
    """The loop variable of :920 / :962 in displacement units."""
    mn, mx, dinc = lo * rng, hi * rng, inc * rng
    out, cur = [], mn
    while cur <= mx + 1e-10:
        out.append(cur)
        cur += dinc
    return out, cur, dinc


def expected(pos0):
    ys = [y for _, y in pos0]
    miny, rng = min(ys), max(ys) - min(ys)
    comp, cur_c, dinc_c = ramp_labels(*CRAMP, rng)
    max_comp = cur_c - dinc_c  # :957
    shear, _, dinc_s = ramp_labels(*SRAMP, rng)
    edata, pos_files = [], []
    s = 1.0
    for k, lab in enumerate(comp):
        s *= 1.0 - CRAMP[2]
        es = stretch_energy(s, 0.0)
        edata.append(COMP_FMT % (0.0, lab, es, 0.0, es))
    x = [p[0] for p in pos0]
    y = [miny + (p[1] - miny) * s for p in pos0]
    for m, lab in enumerate(shear):
        g = (m + 1) * SRAMP[2]
        es = stretch_energy(s, g)
        edata.append(SHEAR_FMT % (lab, max_comp, es, 0.0, es))
        pos_files.append([(x[i] + g * (y[i] - miny), y[i]) for i in range(len(x))])
    digits = len(str(1 + int((SRAMP[1] - SRAMP[0]) / SRAMP[2])))
    return edata, pos_files, digits, len(comp) + len(shear)


def _numbers(line):
    return [float(t) for t in line.split("\t")]


def same_text(got, want_fmt, values, rel=1e-12):
    """got == want_fmt % values, unless a value lies so near a rounding
    boundary of its printed digits that 1e-12 of it changes the text: then
    the text must still have want_fmt's form and the values agree to rel."""
    want = want_fmt % values
    if got == want:
        return True
    lo = want_fmt % tuple(v * (1 - rel) for v in values)
    hi = want_fmt % tuple(v * (1 + rel) for v in values)
    if lo == want and hi == want:
        return False
    if want_fmt % tuple(_numbers(got)) != got:
        return False
    return all(abs(a - b) <= rel * abs(b) + 1e-300 for a, b in zip(_numbers(got), values))


def check_outputs(d, stem, pos0):
    edata_want, pos_want, digits, n_relax = expected(pos0)
    assert digits == 2
    edata = open(os.path.join(d, stem + "_edata.txt")).read().split("\n")
    assert edata[-1] == "" and len(edata) - 1 == len(edata_want) == 14
    for got, want in zip(edata, edata_want):
        fmt = COMP_FMT if got.startswith(" " * 11 + "0\t") else SHEAR_FMT
        assert same_text(got, fmt, tuple(_numbers(want))), (got, want)
        # the label and max_comp columns are exact (no rounding question)
        assert got.split("\t")[:2] == want.split("\t")[:2]
    log = open(os.path.join(d, stem + "_log.txt")).read().split("\n")
    assert log[-1] == "" and len(log) - 1 == n_relax + 1
    for line in log[:n_relax]:
        m = re.fullmatch(r"Convergence reached: (Steps: 1, RMS Force: (.{12,}))", line)
        assert m, line
        rms = float(m.group(2))
        assert rms < 1e-10 and m.group(1) == LOG_FMT % (1, rms)
    assert re.fullmatch(r"Elapsed time: \d+\.\d{6} seconds", log[n_relax])
    for k, pw in enumerate(pos_want, start=1):
        name = os.path.join(d, f"{stem}_{k:02d}.pos")
        lines = open(name).read().split("\n")
        assert lines[-1] == "" and len(lines) - 1 == L * L
        for got, (x, y) in zip(lines, pw):
            assert same_text(got, POS_FMT, (x, y), rel=1e-13), (name, got, (x, y))
    assert not os.path.exists(os.path.join(d, f"{stem}_{len(pos_want) + 1:02d}.pos"))


def test_oracle_protocol_writes_the_golden_files(tmp_path):
    d = str(tmp_path)
    pos0 = write_full_lattice(os.path.join(d, "full.dat"))
    with open(os.path.join(d, "b.txt"), "w") as f:
        f.write(BATCH.format(base=os.path.join(d, "full")))
    rc, base, tag, b = O.parse_batch(os.path.join(d, "b.txt"))
    assert rc == 0 and tag == "g"
    stem = os.path.join(d, "full_g")
    assert O.run_protocol(os.path.join(d, "full.dat"), stem + "_log.txt", stem + "_edata.txt",
                          stem, b) == 0
    check_outputs(d, "full_g", pos0)


def test_a_planted_width_mistake_is_caught(tmp_path):
    """The checker itself: one edata width changed, one label shifted by
    an increment and one .pos precision changed each fail it."""
    d = str(tmp_path)
    pos0 = write_full_lattice(os.path.join(d, "full.dat"))
    with open(os.path.join(d, "b.txt"), "w") as f:
        f.write(BATCH.format(base=os.path.join(d, "full")))
    rc, base, tag, b = O.parse_batch(os.path.join(d, "b.txt"))
    stem = os.path.join(d, "full_g")
    O.run_protocol(os.path.join(d, "full.dat"), stem + "_log.txt", stem + "_edata.txt", stem, b)
    edata_path = stem + "_edata.txt"
    good = open(edata_path).read()
    lines = good.split("\n")
    _, _, _, _ = expected(pos0)
    mutants = []
    # %12.8le -> %12.10le in a compression energy column
    v = _numbers(lines[0])
    mutants.append(("edata", "\n".join([("%12.0lf\t%12.10lf\t%12.10le\t%12.8le\t%12.8le" % tuple(v))]
                                        + lines[1:])))
    # the first shear label without the lag (labels the applied strain)
    v = _numbers(lines[3])
    v[0] += v[0] if v[0] else _numbers(lines[4])[0]
    mutants.append(("edata", "\n".join(lines[:3] + [SHEAR_FMT % tuple(v)] + lines[4:])))
    # max_comp as the applied compression instead of the last label
    v = _numbers(lines[5])
    v[1] += _numbers(lines[1])[1]
    mutants.append(("edata", "\n".join(lines[:5] + [SHEAR_FMT % tuple(v)] + lines[6:])))
    for kind, text in mutants:
        with open(edata_path, "w") as f:
            f.write(text)
        with pytest.raises(AssertionError):
            check_outputs(d, "full_g", pos0)
    with open(edata_path, "w") as f:
        f.write(good)
    check_outputs(d, "full_g", pos0)
    p1 = stem + "_01.pos"
    pl = open(p1).read().split("\n")
    x, y = _numbers(pl[5])
    pl[5] = "%2.9le\t%2.10le" % (x, y)
    with open(p1, "w") as f:
        f.write("\n".join(pl))
    with pytest.raises(AssertionError):
        check_outputs(d, "full_g", pos0)


@pytest.mark.gpu
def test_cli_writes_the_golden_files(tmp_path):
    print("case golden protocol through nrelax", flush=True)
    d = str(tmp_path)
    pos0 = write_full_lattice(os.path.join(d, "full.dat"))
    with open(os.path.join(d, "b.txt"), "w") as f:
        f.write(BATCH.format(base=os.path.join(d, "full")))
    res = subprocess.run([N.CLI_PATH, "-b", os.path.join(d, "b.txt")], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    check_outputs(d, "full_g", pos0)


# ---- the network + gel composite program (network_mesh_relax.cu) ---------
#
# A square network (horizontal bonds, x-periodic, and vertical bonds) on an
# (L + 1) x (M + 1) grid mesh of the strip x0 <= x <= x0 + L, each unit square cut
# into two facets of area 1/2; the column x = L is mesh only, the twin of
# column x0 (read_network_mesh's pairs, :219-226).  After the dilation by
# s0 = sqrt(1 + d) about (min x, min y) (:1254) any affine state is again an
# equilibrium (uniform gel stress, every free node an inversion centre of its
# bonds).  With heights scaled by s_y and shear gamma (relative to the
# current height) from the dilated state, F = [[s0, gamma s_y s0],
# [0, s_y s0]] against the undilated facets, so
#   E_stretch = ks/2 [n_h (s0 - 1)^2 + n_v (s0 s_y sqrt(1 + gamma^2) - 1)^2]
#   E_gel     = A/2 [smod ((1 + s_y^2 (1 + gamma^2)) / s_y - 2)
#                    + bmod (s0^2 s_y - 1 - d_pref)^2]
# with d_pref = d (bmod + net_bmod) / bmod, net_bmod = 2 E_stretch(dilated) /
# (area d^2), area the undilated bounding box (:1217, :1256-1275).  The
# compression ramp starts by applying its minimum and shrinks the height
# (range) by each later step, labels equal to the applied displacement
# (:1290-1328); the shear ramp applies its minimum, then each increment after
# a relax (:1336-1376).  Line formats: :1313, :1318, :1364, :339, :1381.

ML, MM = 6, 5
GEL = dict(ks=1.0, bmod=0.5, smod=0.25, dil=0.05)
MCRAMP = (0.01, 0.03, 0.01)
MSRAMP = (0.0, 0.1, 0.01)
MBATCH = "{base} m\n%r 0 %r %r %r\n0.1 0.99 0.1 1.1 0.5\n5 1.0 1e-10 100000\n%r %r %r\n%r %r %r\n" % (
    (GEL["ks"], GEL["bmod"], GEL["smod"], GEL["dil"]) + MCRAMP + MSRAMP)
X0 = 0.25  # x of the mesh's left column: no free coordinate is exactly 0
MCOMP_FMT = "%12.0lf\t%12.10lf\t%12.8le\t%12.8le\t%12.8le\t%12.8le"  # :1318
MSHEAR_FMT = "%12.8lf\t%12.8lf\t%12.8le\t%12.8le\t%12.8le\t%12.8le"  # :1364


def write_grid_composite(path):
    pts = [(X0 + c, float(r)) for r in range(MM + 1) for c in range(ML + 1)]
    idx = lambda r, c: r * (ML + 1) + c  # noqa: E731
    edges = []
    for r in range(MM + 1):
        for c in range(ML):
            edges.append((idx(r, c), idx(r, (c + 1) % ML), 1 if c == ML - 1 else 0))
            if r < MM:
                edges.append((idx(r, c), idx(r + 1, c), 0))
    facets = []
    for r in range(MM):
        for c in range(ML):
            a, b, cc, d = idx(r, c), idx(r, c + 1), idx(r + 1, c + 1), idx(r + 1, c)
            facets += [(a, b, cc), (a, cc, d)]
    with open(path, "w") as f:
        f.write(f"{len(pts)} {len(edges)} {len(facets)} {float(ML)!r}\n")
        for x, y in pts:
            f.write(f"{x!r} {y!r}\n")
        for i, j, w in edges:
            f.write(f"{i} {j} {w}\n")
        for t in facets:
            f.write("%d %d %d\n" % t)
        for _ in facets:
            f.write("0.5\n")
    return pts


def composite_expected(pts):
    ks, bmod, smod, d = GEL["ks"], GEL["bmod"], GEL["smod"], GEL["dil"]
    s0 = math.sqrt(1.0 + d)
    n_h, n_v = ML * (MM + 1), ML * MM
    area_box = float(ML) * float(MM)
    a_tot = 0.5 * 2 * ML * MM
    net_bmod = 2.0 * (0.5 * ks * (n_h + n_v) * (s0 - 1.0) ** 2) / area_box / (d * d)
    pref = d * (bmod + net_bmod) / bmod

    def energies(s_y, g):
        es = 0.5 * ks * (n_h * (s0 - 1.0) ** 2 +
                         n_v * (s0 * s_y * math.sqrt(1.0 + g * g) - 1.0) ** 2)
        eg = 0.5 * a_tot * (smod * ((1.0 + s_y * s_y * (1.0 + g * g)) / s_y - 2.0) +
                            bmod * (s0 * s0 * s_y - 1.0 - pref) ** 2)
        return es, eg

    range0 = float(MM) * s0
    rng = range0
    mn, mx, inc = (v * range0 for v in MCRAMP)
    lines, pos_files = [], []
    rng -= mn
    cur = mn
    while cur <= mx + 1e-10:
        es, eg = energies(rng / range0, 0.0)
        lines.append(MCOMP_FMT % (0.0, cur, es, 0.0, eg, es + eg))
        if cur < mx:
            rng -= inc
        cur += inc
    max_comp = cur - inc if mx > mn else mx
    s_y = rng / range0
    mn, mx, inc = (v * range0 for v in MSRAMP)
    applied, cur = mn, mn
    while cur <= mx + 1e-10:
        g = applied / rng
        es, eg = energies(s_y, g)
        lines.append(MSHEAR_FMT % (cur, max_comp, es, 0.0, eg, es + eg))
        pos_files.append([(X0 + s0 * (x - X0) + g * s_y * s0 * y, s_y * s0 * y) for x, y in pts])
        if cur < mx:
            applied += inc
        cur += inc
    return lines, pos_files, pref


def check_composite_outputs(d, stem, pts):
    want, pos_want, _ = composite_expected(pts)
    assert len(want) == 3 + 11 and len(pos_want) == 11
    got = open(os.path.join(d, stem + "_edata.txt")).read().split("\n")
    assert got[-1] == "" and len(got) - 1 == len(want)
    for k, (g, w) in enumerate(zip(got, want)):
        fmt = MCOMP_FMT if k < 3 else MSHEAR_FMT
        assert same_text(g, fmt, tuple(_numbers(w)), rel=1e-11), (g, w)
        assert g.split("\t")[:2] == w.split("\t")[:2]
    log = open(os.path.join(d, stem + "_log.txt")).read().split("\n")
    assert log[-1] == "" and len(log) - 1 == len(want) + 1
    for line in log[:len(want)]:
        m = re.fullmatch(r"Convergence reached: (Steps: 1, RMS Force: (.{12,}))", line)
        assert m, line
        assert float(m.group(2)) < 1e-10 and m.group(1) == LOG_FMT % (1, float(m.group(2)))
    assert re.fullmatch(r"Elapsed time: \d+\.\d{6} seconds", log[len(want)])
    for k, pw in enumerate(pos_want, start=1):
        lines = open(os.path.join(d, f"{stem}_{k:02d}.pos")).read().split("\n")
        assert lines[-1] == "" and len(lines) - 1 == len(pts)
        for g, (x, y) in zip(lines, pw):
            assert same_text(g, POS_FMT, (x, y), rel=1e-13), (k, g, (x, y))


def test_oracle_composite_protocol_writes_the_golden_files(tmp_path):
    d = str(tmp_path)
    pts = write_grid_composite(os.path.join(d, "grid.nmc"))
    with open(os.path.join(d, "b.txt"), "w") as f:
        f.write(MBATCH.format(base=os.path.join(d, "grid")))
    rc, base, tag, b = O.parse_mesh_batch(os.path.join(d, "b.txt"))
    assert rc == 0 and tag == "m"
    stem = os.path.join(d, "grid_m")
    rc, pd = O.run_mesh_protocol(os.path.join(d, "grid.nmc"), stem + "_log.txt",
                                 stem + "_edata.txt", stem, b)
    assert rc == 0
    assert abs(pd - composite_expected(pts)[2]) <= 1e-14 * pd
    check_composite_outputs(d, "grid_m", pts)


@pytest.mark.gpu
def test_mesh_cli_writes_the_golden_files(tmp_path):
    print("case golden composite protocol through nrelax_mesh", flush=True)
    d = str(tmp_path)
    pts = write_grid_composite(os.path.join(d, "grid.nmc"))
    with open(os.path.join(d, "b.txt"), "w") as f:
        f.write(MBATCH.format(base=os.path.join(d, "grid")))
    res = subprocess.run([N.MESH_CLI_PATH, "-b", os.path.join(d, "b.txt")], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    # cout's default precision (6 significant digits) for the preferred dilation
    assert "Preferred dilation: %s" % ("%.6g" % composite_expected(pts)[2]) in res.stdout
    check_composite_outputs(d, "grid_m", pts)
