// Host-side network model.  Compiled without floating-point contraction so
// that rest lengths and the collinearity test round exactly as the
// reference's host code does.
#include "network.hpp"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <numeric>

#include "../../include/nrelax.h"

namespace nr {

namespace {

constexpr double kFltTol = 1e-10;  // FLT_TOL, large_strain_relax.cu:40

// Image offset of a wrapped neighbour at read time (no Lees-Edwards shift
// yet): (wx*W, wy*H).  In reference mode wy is 0 and the y term is +0.0.
inline void bond_vector0(const std::vector<double> &p, double W, double H,
                         int i, int j, int wx, int wy, double &dx, double &dy)
{
    dx = p[2 * j] + W * wx - p[2 * i];
    dy = p[2 * j + 1] + H * wy - p[2 * i + 1];
}

inline double rest_length(const std::vector<double> &p, double W, double H,
                          int i, int j, int wx, int wy)
{
    double dx, dy;
    bond_vector0(p, W, H, i, j, wx, wy, dx, dy);
    double a = dx * dx;
    double b = dy * dy;
    return std::sqrt(a + b);
}

uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

double uniform01(uint64_t seed, uint64_t counter)
{
    uint64_t h = mix64(mix64(seed) + counter);
    return (double)(h >> 11) * 0x1.0p-53;
}

// read_network, large_strain_relax.cu:42-114, with std::getline/eof()
// semantics kept (a last point line without a newline reads as "too few
// points", :81-85).  Extensions: 4th header number = box height (Lees-Edwards
// mode); in that mode a 4th edge integer is the y wrap.
int read_network(const std::string &path, Network &net, std::string &err)
{
    std::ifstream in(path);
    std::string line;
    net = Network();
    if (!in.is_open()) {
        err = "cannot open " + path;
        return NR_E_OPEN;
    }
    int n = 0, e = 0;
    double W = 0.0, H = 0.0;
    std::getline(in, line);
    int nr = std::sscanf(line.c_str(), "%d %d %lf %lf", &n, &e, &W, &H);
    if (nr < 3 || W < 0 || n <= 0 || e < 0) {
        err = "Reading failed. The provided file has an invalid header.";
        return NR_E_HEADER;
    }
    if (nr < 4 || H < 0) H = 0.0;
    net.n = n;
    net.width = W;
    net.height = H;
    double miny = FLT_MAX, maxy = FLT_MIN;
    net.pos.reserve(2 * (size_t)n);
    while (!in.eof() && net.pos.size() < 2 * (size_t)n) {
        std::getline(in, line);
        double x, y;
        if (std::sscanf(line.c_str(), "%lf %lf", &x, &y) == 2) {
            net.pos.push_back(x);
            net.pos.push_back(y);
            if (y < miny) miny = y;
            if (y > maxy) maxy = y;
        }
    }
    if (in.eof()) {
        err = "Reading failed. Too few points were read.";
        return NR_E_POINTS;
    }
    net.miny = miny;
    net.maxy = maxy;
    while (!in.eof() && (int)net.rest.size() < e) {
        std::getline(in, line);
        int i1, i2;
        signed char wx, wy;
        nr = std::sscanf(line.c_str(), "%d %d %hhd %hhd", &i1, &i2, &wx, &wy);
        if (nr < 3) continue;
        if (std::min(i1, i2) < 0 || std::max(i1, i2) >= n) {
            err = "Reading failed due to an out-of-bounds point index.";
            return NR_E_INDEX;
        }
        int wyi = (H > 0 && nr == 4) ? wy : 0;
        net.bonds.push_back(i1);
        net.bonds.push_back(i2);
        net.bwrap.push_back(wx);
        net.bwrap.push_back((signed char)wyi);
        net.rest.push_back(rest_length(net.pos, W, H, i1, i2, wx, wyi));
    }
    if ((int)net.rest.size() < e) {
        err = "Too few edges were found.";
        return NR_E_EDGES;
    }
    set_clamps(net);
    return NR_OK;
}

int network_from_arrays(int n, int e, double width, double height,
                        const double *pos, const int *bonds,
                        const signed char *wraps, Network &net,
                        std::string &err)
{
    net = Network();
    if (n <= 0 || e < 0 || width < 0 || height < 0 || !pos || (e > 0 && (!bonds || !wraps))) {
        err = "invalid network arrays";
        return NR_E_ARG;
    }
    net.n = n;
    net.width = width;
    net.height = height;
    net.pos.assign(pos, pos + 2 * (size_t)n);
    double miny = FLT_MAX, maxy = FLT_MIN;
    for (int i = 0; i < n; i++) {
        double y = pos[2 * i + 1];
        if (y < miny) miny = y;
        if (y > maxy) maxy = y;
    }
    net.miny = miny;
    net.maxy = maxy;
    for (int b = 0; b < e; b++) {
        int i1 = bonds[2 * b], i2 = bonds[2 * b + 1];
        if (std::min(i1, i2) < 0 || std::max(i1, i2) >= n) {
            err = "bond index out of range";
            return NR_E_INDEX;
        }
        int wx = wraps[2 * b], wy = height > 0 ? wraps[2 * b + 1] : 0;
        net.bonds.push_back(i1);
        net.bonds.push_back(i2);
        net.bwrap.push_back((signed char)wx);
        net.bwrap.push_back((signed char)wy);
        net.rest.push_back(rest_length(net.pos, width, height, i1, i2, wx, wy));
    }
    set_clamps(net);
    return NR_OK;
}

// Clamp codes, large_strain_relax.cu:896-902 (reference mode only).
void set_clamps(Network &net)
{
    net.fcode.assign(net.n, 3);
    if (net.height > 0) return;
    for (int i = 0; i < net.n; i++) {
        double y = net.pos[2 * i + 1];
        if (y < net.miny + kFltTol || y > net.maxy - kFltTol) net.fcode[i] = 0;
    }
}

// get_triples, large_strain_relax.cu:116-184: per-node neighbour lists in
// edge-file order (reverse direction with negated wrap), hinges ascending,
// pairs (i < j) in list order, antiparallel test in the input geometry.
int build_triples(Network &net)
{
    struct Adj {
        int nbr, bond;
        signed char wx, wy;
    };
    std::vector<std::vector<Adj>> adj(net.n);
    for (int b = 0; b < net.e(); b++) {
        int i = net.bonds[2 * b], j = net.bonds[2 * b + 1];
        signed char wx = net.bwrap[2 * b], wy = net.bwrap[2 * b + 1];
        adj[i].push_back({j, b, wx, wy});
        adj[j].push_back({i, b, (signed char)-wx, (signed char)-wy});
    }
    net.tri.clear();
    net.twrap.clear();
    net.tbond.clear();
    net.tkb.clear();
    const double W = net.width, H = net.height;
    for (int h = 0; h < net.n; h++) {
        const std::vector<Adj> &nb = adj[h];
        for (size_t i = 0; i < nb.size(); i++) {
            double dx2, dy2, dx3, dy3;
            bond_vector0(net.pos, W, H, h, nb[i].nbr, nb[i].wx, nb[i].wy, dx2, dy2);
            for (size_t j = i + 1; j < nb.size(); j++) {
                bond_vector0(net.pos, W, H, h, nb[j].nbr, nb[j].wx, nb[j].wy, dx3, dy3);
                double p1 = dx2 * dx3, p2 = dy2 * dy3;
                double dot = p1 + p2;
                double m1 = dx2 * dx2, m2 = dy2 * dy2, m3 = dx3 * dx3, m4 = dy3 * dy3;
                double mag_sq = (m1 + m2) * (m3 + m4);
                double dd = dot * dot;
                if (dot < 0 && std::fabs(dd - mag_sq) < kFltTol) {
                    net.tri.push_back(h);
                    net.tri.push_back(nb[i].nbr);
                    net.tri.push_back(nb[j].nbr);
                    net.twrap.push_back(nb[i].wx);
                    net.twrap.push_back(nb[i].wy);
                    net.twrap.push_back(nb[j].wx);
                    net.twrap.push_back(nb[j].wy);
                    net.tbond.push_back(nb[i].bond);
                    net.tbond.push_back(nb[j].bond);
                }
            }
        }
    }
    net.triples_built = true;
    return NR_OK;
}

// Fibers: bonds joined through triples form one fiber; the label is the
// smallest bond index of the fiber.
int fiber_labels(const Network &net, std::vector<int> &label)
{
    std::vector<int> parent(net.e());
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int a) {
        while (parent[a] != a) {
            parent[a] = parent[parent[a]];
            a = parent[a];
        }
        return a;
    };
    for (int t = 0; t < net.t(); t++) {
        int a = find(net.tbond[2 * t]), b = find(net.tbond[2 * t + 1]);
        if (a < b) parent[b] = a;
        else if (b < a) parent[a] = b;
    }
    label.resize(net.t());
    for (int t = 0; t < net.t(); t++) label[t] = find(net.tbond[2 * t]);
    return NR_OK;
}

// kappa = exp(ln lo + u (ln hi - ln lo)), u uniform from (seed, fiber label).
int set_fiber_kappa(Network &net, double lo, double hi, uint64_t seed)
{
    if (!(lo > 0) || !(hi >= lo)) return NR_E_ARG;
    if (!net.triples_built) return NR_E_ARG;
    std::vector<int> label;
    fiber_labels(net, label);
    double llo = std::log(lo), lhi = std::log(hi);
    net.tkb.resize(net.t());
    for (int t = 0; t < net.t(); t++) {
        double u = uniform01(seed, (uint64_t)label[t]);
        net.tkb[t] = std::exp(llo + u * (lhi - llo));
    }
    return NR_OK;
}

// Triangular lattice: node r*L + c at (c + (r odd)/2, r*sqrt(3)/2); candidate
// bonds per node in directions 0, 60 and 120 degrees (candidate id 3*node+d),
// periodic in x; in Lees-Edwards mode also periodic in y (L even).
int make_lattice(int L, double p, uint64_t seed, bool le, Network &net,
                 std::string &err)
{
    if (L < 3 || (le && (L & 1))) {
        err = "lattice size must be >= 3 (even in Lees-Edwards mode)";
        return NR_E_ARG;
    }
    const int n = L * L;
    const double hgt = std::sqrt(3.0) * 0.5;
    std::vector<double> pos(2 * (size_t)n);
    std::vector<int> bonds;
    std::vector<signed char> wraps;
    for (int r = 0; r < L; r++)
        for (int c = 0; c < L; c++) {
            pos[2 * (r * L + c)] = c + ((r & 1) ? 0.5 : 0.0);
            pos[2 * (r * L + c) + 1] = r * hgt;
        }
    for (int r = 0; r < L; r++)
        for (int c = 0; c < L; c++)
            for (int d = 0; d < 3; d++) {
                int node = r * L + c, r2 = r, c2, wx = 0, wy = 0;
                if (d == 0) {
                    c2 = c + 1;
                } else {
                    r2 = r + 1;
                    c2 = (d == 1) ? c + (r & 1) : c - 1 + (r & 1);
                }
                if (c2 == L) { c2 = 0; wx = 1; }
                if (c2 < 0) { c2 = L - 1; wx = -1; }
                if (r2 == L) {
                    if (!le) continue;
                    r2 = 0;
                    wy = 1;
                }
                if (!(uniform01(seed, (uint64_t)node * 3 + (uint64_t)d) < p)) continue;
                bonds.push_back(node);
                bonds.push_back(r2 * L + c2);
                wraps.push_back((signed char)wx);
                wraps.push_back((signed char)wy);
            }
    return network_from_arrays(n, (int)bonds.size() / 2, (double)L, le ? L * hgt : 0.0,
                               pos.data(), bonds.data(), wraps.data(), net, err);
}

// Writes a .dat the reference can read (3-number header in reference mode,
// positions with round-trip precision).
int write_dat(const Network &net, const std::string &path)
{
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) return NR_E_OPEN;
    if (net.height > 0)
        std::fprintf(fp, "%d %d %.17g %.17g\n", net.n, net.e(), net.width, net.height);
    else
        std::fprintf(fp, "%d %d %.17g\n", net.n, net.e(), net.width);
    for (int i = 0; i < net.n; i++)
        std::fprintf(fp, "%.17g %.17g\n", net.pos[2 * i], net.pos[2 * i + 1]);
    for (int b = 0; b < net.e(); b++) {
        if (net.height > 0)
            std::fprintf(fp, "%d %d %d %d\n", net.bonds[2 * b], net.bonds[2 * b + 1],
                         (int)net.bwrap[2 * b], (int)net.bwrap[2 * b + 1]);
        else
            std::fprintf(fp, "%d %d %d\n", net.bonds[2 * b], net.bonds[2 * b + 1],
                         (int)net.bwrap[2 * b]);
    }
    std::fclose(fp);
    return NR_OK;
}

}  // namespace nr
