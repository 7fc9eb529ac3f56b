// Host-side network model: the reference's .dat reader, bending-triple
// enumeration, clamp codes, fiber labels and the synthetic lattice.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace nr {

struct Network {
    int n = 0;
    double width = 0.0;   // x period
    double height = 0.0;  // y period; 0 = reference mode (clamped rows)
    double miny = 0.0, maxy = 0.0;
    std::vector<double> pos;          // 2n input geometry
    std::vector<int> bonds;           // 2e
    std::vector<signed char> bwrap;   // 2e (wx, wy)
    std::vector<double> rest;         // e
    std::vector<int> tri;             // 3t (hinge, end a, end b)
    std::vector<signed char> twrap;   // 4t (wx_a, wy_a, wx_b, wy_b)
    std::vector<int> tbond;           // 2t bonds of each triple
    std::vector<double> tkb;          // t per-triple kappa, or empty
    std::vector<unsigned char> fcode; // n clamp codes
    bool triples_built = false;

    int e() const { return (int)rest.size(); }
    int t() const { return (int)tri.size() / 3; }
};

// Status codes match include/nrelax.h.
int read_network(const std::string &path, Network &net, std::string &err);
int network_from_arrays(int n, int e, double width, double height,
                        const double *pos, const int *bonds,
                        const signed char *wraps, Network &net,
                        std::string &err);
void set_clamps(Network &net);
int build_triples(Network &net);
int fiber_labels(const Network &net, std::vector<int> &label);
int set_fiber_kappa(Network &net, double lo, double hi, uint64_t seed);
int make_lattice(int L, double p, uint64_t seed, bool lees_edwards,
                 Network &net, std::string &err);
int write_dat(const Network &net, const std::string &path);

double uniform01(uint64_t seed, uint64_t counter);

}  // namespace nr
