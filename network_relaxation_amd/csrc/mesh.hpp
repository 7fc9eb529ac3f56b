// Host-side model of a network + FEM-gel composite (the reference's
// src/network_mesh_relax.cu): a bond/bend network on the points of a
// triangular mesh whose facets are a compressible neo-Hookean gel, periodic in
// x through twin points (each left-edge point has a twin one box width to the
// right; the two receive the sum of their forces).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "network.hpp"

namespace nr {

struct Mesh {
    Network net;                // all points, bonds, clamp codes (triples at set-up)
    std::vector<int> facets;    // 3 per facet
    std::vector<double> areas;  // per facet, as read
    std::vector<int> twins;     // pairs (left-edge point, its twin)
    double offset = 0.0;        // header offset as read
    double minx = 0.0, maxx = 0.0;

    int nf() const { return (int)areas.size(); }
    int np() const { return (int)twins.size() / 2; }
};

// The composite after the reference's set-up for one dilation
// (network_mesh_relax.cu:1138-1275).
struct MeshSetup {
    std::vector<double> pos;  // 2n dilated points
    std::vector<double> inv;  // 6 per facet, from the undilated points
    double width = 0.0;       // box offset x sqrt(1 + dilation)
    double range = 0.0;       // (maxy - miny) x sqrt(1 + dilation)
    double pref_d = 0.0;      // the gel's preferred dilation
    double net_bmod = 0.0;    // the network's bulk modulus estimate
};

// Status codes match include/nrelax.h.
int read_mesh(const std::string &path, Mesh &m, std::string &err);
int make_mesh_lattice(int L, double p, uint64_t seed, Mesh &m, std::string &err);
int write_nmc(const Mesh &m, const std::string &path);
// Triples (kb > 0, found with the dilated offset), inverse matrices, dilated
// start and preferred dilation.  Changes m.net (width, triples).
int mesh_setup(Mesh &m, double ks, double kb, double bmod, double dilation, MeshSetup &s,
               std::string &err);

}  // namespace nr
