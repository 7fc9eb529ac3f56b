// Host-side composite model: the reference's .nmc reader
// (network_mesh_relax.cu:72-230), twin pairs (:219-226), the set-up of
// do_relaxation_run (:1138-1275) and a synthetic composite lattice.
// Compiled without floating-point contraction, like network.cpp.
#include "mesh.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <fstream>

#include "../../include/nrelax.h"

namespace nr {

namespace {

constexpr double kFltTol = 1e-10;  // FLT_TOL, network_mesh_relax.cu:40

bool coincident(const std::vector<double> &p, int a, double x, double y)
{
    return std::fabs(p[2 * a] - x) <= kFltTol && std::fabs(p[2 * a + 1] - y) <= kFltTol;
}

// Twins (:219-226): each left-edge point (x < min_x + 1e-10; the first of
// coincident ones), in x-then-y order with the reference's 1e-10 tolerance,
// paired with the first point within 1e-10 of (x + offset, y).  Only points
// near the right edge can match, so both sides are small lists.
void find_twins(Mesh &m)
{
    const std::vector<double> &p = m.net.pos;
    std::vector<int> left, right;
    for (int i = 0; i < m.net.n; i++) {
        double x = p[2 * i];
        if (x < m.minx + kFltTol) left.push_back(i);
        if (x >= m.minx + m.offset - 2 * kFltTol && x <= m.minx + m.offset + 3 * kFltTol)
            right.push_back(i);
    }
    std::sort(left.begin(), left.end(), [&](int a, int b) {
        if (p[2 * a] < p[2 * b] - kFltTol) return true;
        if (p[2 * a] > p[2 * b] + kFltTol) return false;
        if (p[2 * a + 1] < p[2 * b + 1] - kFltTol) return true;
        if (p[2 * a + 1] > p[2 * b + 1] + kFltTol) return false;
        return a < b;
    });
    m.twins.clear();
    for (int a : left) {
        bool dup = false;
        for (int j : left)
            if (j < a && coincident(p, j, p[2 * a], p[2 * a + 1])) {
                dup = true;
                break;
            }
        if (dup) continue;
        int q = -1;
        for (int j : right)
            if (coincident(p, j, p[2 * a] + m.offset, p[2 * a + 1]) && (q < 0 || j < q)) q = j;
        if (q >= 0) {
            m.twins.push_back(a);
            m.twins.push_back(q);
        }
    }
}

}  // namespace

// read_network_mesh :72-230 with std::getline/eof() semantics: header
// "%d %d %d %lf" (points, edges, facets, offset); each section takes the
// lines that parse; facets must be followed by one positive area each.
int read_mesh(const std::string &path, Mesh &m, std::string &err)
{
    std::ifstream in(path);
    std::string line;
    m = Mesh();
    Network &net = m.net;
    if (!in.is_open()) {
        err = "cannot open " + path;
        return NR_E_OPEN;
    }
    int n = 0, e = 0, nf = 0;
    double W = 0.0;
    std::getline(in, line);
    int nr = std::sscanf(line.c_str(), "%d %d %d %lf", &n, &e, &nf, &W);
    if (nr < 4 || W < 0 || n <= 0 || e < 0 || nf < 0) {
        err = "Reading failed. The provided file has an invalid header.";
        return NR_E_HEADER;
    }
    net.n = n;
    net.width = W;
    m.offset = W;
    double minx = FLT_MAX, miny = FLT_MAX, maxx = FLT_MIN, maxy = FLT_MIN;
    net.pos.reserve(2 * (size_t)n);
    while (!in.eof() && net.pos.size() < 2 * (size_t)n) {
        std::getline(in, line);
        double x, y;
        if (std::sscanf(line.c_str(), "%lf %lf", &x, &y) == 2) {
            net.pos.push_back(x);
            net.pos.push_back(y);
            minx = std::min(minx, x);
            miny = std::min(miny, y);
            maxx = std::max(maxx, x);
            maxy = std::max(maxy, y);
        }
    }
    if (in.eof()) {
        err = "Reading failed. Too few points were read.";
        return NR_E_POINTS;
    }
    net.miny = miny;
    net.maxy = maxy;
    m.minx = minx;
    m.maxx = maxx;
    while (!in.eof() && (int)net.rest.size() < e) {
        std::getline(in, line);
        int i1, i2;
        signed char wx;
        if (std::sscanf(line.c_str(), "%d %d %hhd", &i1, &i2, &wx) != 3) continue;
        if (std::min(i1, i2) < 0 || std::max(i1, i2) >= n) {
            err = "Reading failed due to an out-of-bounds point index.";
            return NR_E_INDEX;
        }
        double dx = net.pos[2 * i2] + W * wx - net.pos[2 * i1];
        double dy = net.pos[2 * i2 + 1] - net.pos[2 * i1 + 1];
        double a = dx * dx, b = dy * dy;
        net.bonds.push_back(i1);
        net.bonds.push_back(i2);
        net.bwrap.push_back(wx);
        net.bwrap.push_back(0);
        net.rest.push_back(std::sqrt(a + b));
    }
    if ((int)net.rest.size() < e) {
        err = "Too few edges were read.";
        return NR_E_EDGES;
    }
    if (in.eof() && nf > 0) {
        err = "The file ended before facets could be read.";
        return NR_E_MESH;
    }
    while (!in.eof() && (int)m.facets.size() < 3 * nf) {
        std::getline(in, line);
        int a, b, c;
        if (std::sscanf(line.c_str(), "%d %d %d", &a, &b, &c) != 3) continue;
        if (a < 0 || a >= n || b < 0 || b >= n || c < 0 || c >= n) {
            err = "Reading failed due to an out-of-bounds point index.";
            return NR_E_INDEX;
        }
        m.facets.push_back(a);
        m.facets.push_back(b);
        m.facets.push_back(c);
    }
    if ((int)m.facets.size() < 3 * nf) {
        err = "Too few facets were read.";
        return NR_E_MESH;
    }
    if (in.eof() && nf > 0) {
        err = "The file ended before facet areas could be read.";
        return NR_E_MESH;
    }
    while (!in.eof() && (int)m.areas.size() < nf) {
        std::getline(in, line);
        double a;
        if (std::sscanf(line.c_str(), "%lf", &a) != 1) continue;
        if (!(a > 0)) {
            err = "An illegal area was specified.";
            return NR_E_MESH;
        }
        m.areas.push_back(a);
    }
    if ((int)m.areas.size() < nf) {
        err = "Too few areas were specified.";
        return NR_E_MESH;
    }
    set_clamps(net);
    find_twins(m);
    return NR_OK;
}

// Composite lattice: the reference-mode triangular lattice of make_lattice
// (L x L points, x-periodic bonds kept with probability p, clamped top and
// bottom rows) on a gel mesh of the strip 0 <= x <= L.  Points L*L.. are
// mesh-only: per even row its left point's twin at x = L; per odd row a left
// edge point at x = 0 and its twin at x = L.  Facets: every lattice triangle
// of the strip, the two boundary half-cells of each row pair as right
// triangles; areas are the facets' own.
int make_mesh_lattice(int L, double p, uint64_t seed, Mesh &m, std::string &err)
{
    m = Mesh();
    int rc = make_lattice(L, p, seed, false, m.net, err);
    if (rc != NR_OK) return rc;
    Network &net = m.net;
    const double hgt = std::sqrt(3.0) * 0.5;
    std::vector<int> twin_even(L, -1), left_odd(L, -1), right_odd(L, -1);
    int n = L * L;
    for (int r = 0; r < L; r++) {
        double y = r * hgt;
        if (!(r & 1)) {
            twin_even[r] = n++;
            net.pos.push_back((double)L);
            net.pos.push_back(y);
        } else {
            left_odd[r] = n++;
            net.pos.push_back(0.0);
            net.pos.push_back(y);
            right_odd[r] = n++;
            net.pos.push_back((double)L);
            net.pos.push_back(y);
        }
    }
    net.n = n;
    // point of row r at lattice column c (c == L: the even row's twin)
    auto P = [&](int r, int c) { return c == L ? twin_even[r] : r * L + c; };
    auto tri = [&](int a, int b, int c) {
        m.facets.push_back(a);
        m.facets.push_back(b);
        m.facets.push_back(c);
    };
    for (int r = 0; r + 1 < L; r++) {
        if (!(r & 1)) {  // row r at integer x, row r + 1 at half-integer x
            for (int c = 0; c < L; c++) tri(P(r, c), P(r, c + 1), P(r + 1, c));
            for (int c = 0; c + 1 < L; c++) tri(P(r + 1, c + 1), P(r + 1, c), P(r, c + 1));
            tri(P(r, 0), P(r + 1, 0), left_odd[r + 1]);
            tri(P(r, L), right_odd[r + 1], P(r + 1, L - 1));
        } else {  // row r at half-integer x, row r + 1 at integer x
            for (int c = 0; c + 1 < L; c++) tri(P(r, c), P(r, c + 1), P(r + 1, c + 1));
            for (int c = 0; c < L; c++) tri(P(r + 1, c + 1), P(r + 1, c), P(r, c));
            tri(left_odd[r], P(r, 0), P(r + 1, 0));
            tri(right_odd[r], P(r + 1, L), P(r, L - 1));
        }
    }
    const int nf = (int)m.facets.size() / 3;
    for (int f = 0; f < nf; f++) {
        const int *t = &m.facets[3 * f];
        double x1 = net.pos[2 * t[0]], y1 = net.pos[2 * t[0] + 1];
        double x2 = net.pos[2 * t[1]], y2 = net.pos[2 * t[1] + 1];
        double x3 = net.pos[2 * t[2]], y3 = net.pos[2 * t[2] + 1];
        double det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
        m.areas.push_back(0.5 * std::fabs(det));
    }
    m.offset = (double)L;
    m.minx = 0.0;
    m.maxx = (double)L;
    set_clamps(net);
    find_twins(m);
    return NR_OK;
}

int write_nmc(const Mesh &m, const std::string &path)
{
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) return NR_E_OPEN;
    const Network &net = m.net;
    std::fprintf(fp, "%d %d %d %.17g\n", net.n, net.e(), m.nf(), m.offset);
    for (int i = 0; i < net.n; i++)
        std::fprintf(fp, "%.17g %.17g\n", net.pos[2 * i], net.pos[2 * i + 1]);
    for (int b = 0; b < net.e(); b++)
        std::fprintf(fp, "%d %d %d\n", net.bonds[2 * b], net.bonds[2 * b + 1],
                     (int)net.bwrap[2 * b]);
    for (int f = 0; f < m.nf(); f++)
        std::fprintf(fp, "%d %d %d\n", m.facets[3 * f], m.facets[3 * f + 1], m.facets[3 * f + 2]);
    for (int f = 0; f < m.nf(); f++) std::fprintf(fp, "%.17g\n", m.areas[f]);
    std::fclose(fp);
    return NR_OK;
}

// Set-up of do_relaxation_run :1138-1275.  The box offset grows by
// s = sqrt(1 + dilation) (:1141) before the triples are found on the
// undilated points (:1143-1145), so with dilation > 0 inclined fibers that
// cross the periodic boundary have no bending triples, as in the reference.
// Inverse matrices (get_inv_mats :503-531) come from the undilated points
// (:1250); the points are dilated about (min_x, min_y) (dilate :796-806,
// :1254); the network's stretching energy there gives its bulk modulus
// 2 E / area / dilation^2 (:1256-1271) and the gel's preferred dilation
// dilation (bmod + net_bmod) / bmod (:1274).  With dilation 0 the preferred
// dilation is 0 (the reference divides 0 by 0 there).
int mesh_setup(Mesh &m, double ks, double kb, double bmod, double dilation, MeshSetup &s,
               std::string &err)
{
    if (!(ks >= 0) || !(kb >= 0) || !(bmod >= 0) || !(dilation >= 0)) {
        err = "moduli and dilation must be >= 0";
        return NR_E_ARG;
    }
    Network &net = m.net;
    const double sc = std::sqrt(1.0 + dilation);
    net.width = m.offset * sc;
    net.tri.clear();
    net.twrap.clear();
    net.tbond.clear();
    net.triples_built = false;
    if (kb > 0) {
        int rc = build_triples(net);
        if (rc != NR_OK) {
            err = "triple enumeration failed";
            return rc;
        }
    }
    const int nf = m.nf();
    s.inv.resize(6 * (size_t)nf);
    for (int f = 0; f < nf; f++) {
        const int *t = &m.facets[3 * f];
        double x1 = net.pos[2 * t[0]], y1 = net.pos[2 * t[0] + 1];
        double x2 = net.pos[2 * t[1]], y2 = net.pos[2 * t[1] + 1];
        double x3 = net.pos[2 * t[2]], y3 = net.pos[2 * t[2] + 1];
        double det = std::fma(x1, y2 - y3, std::fma(x2, y3 - y1, x3 * (y1 - y2)));
        double *M = &s.inv[6 * (size_t)f];
        M[0] = (y2 - y3) / det;
        M[1] = (x3 - x2) / det;
        M[2] = (y3 - y1) / det;
        M[3] = (x1 - x3) / det;
        M[4] = (y1 - y2) / det;
        M[5] = (x2 - x1) / det;
    }
    s.pos.resize(2 * (size_t)net.n);
    for (int i = 0; i < net.n; i++) {
        s.pos[2 * i] = std::fma(sc, net.pos[2 * i] - m.minx, m.minx);
        s.pos[2 * i + 1] = std::fma(sc, net.pos[2 * i + 1] - net.miny, net.miny);
    }
    s.width = net.width;
    s.range = (net.maxy - net.miny) * sc;
    s.net_bmod = 0.0;
    if (ks > 0 && dilation > 0) {
        const double area = (m.maxx - m.minx) * (net.maxy - net.miny);
        double es = 0.0;
        for (int b = 0; b < net.e(); b++) {
            int p = net.bonds[2 * b], q = net.bonds[2 * b + 1];
            double rx = s.pos[2 * q] + net.width * net.bwrap[2 * b] - s.pos[2 * p];
            double ry = s.pos[2 * q + 1] - s.pos[2 * p + 1];
            double d = std::sqrt(std::fma(rx, rx, ry * ry));
            double dl = d - net.rest[b];
            es += 0.5 * ks * dl * dl;
        }
        s.net_bmod = es * (2.0 / area / (dilation * dilation));
    }
    s.pref_d = bmod > 0 ? dilation * (bmod + s.net_bmod) / bmod : dilation;
    return NR_OK;
}

}  // namespace nr
