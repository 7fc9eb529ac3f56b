// Device layout builder: Morton-ordered spatial tiles, halos, and per-node
// incidence lists in element order (bonds by index, then triples by index),
// so that a node's force is summed in the same order as a sequential pass over
// the reference's bond and triple kernels.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "../../include/nrelax.h"

namespace nr {

namespace {

uint32_t spread16(uint32_t v)
{
    v &= 0xFFFF;
    v = (v | (v << 8)) & 0x00FF00FF;
    v = (v | (v << 4)) & 0x0F0F0F0F;
    v = (v | (v << 2)) & 0x33333333;
    v = (v | (v << 1)) & 0x55555555;
    return v;
}

// Morton order of the input geometry (ties broken by node index).
std::vector<int> spatial_order(const Network &net)
{
    double minx = 1e300, maxx = -1e300, miny = 1e300, maxy = -1e300;
    for (int i = 0; i < net.n; i++) {
        minx = std::min(minx, net.pos[2 * i]);
        maxx = std::max(maxx, net.pos[2 * i]);
        miny = std::min(miny, net.pos[2 * i + 1]);
        maxy = std::max(maxy, net.pos[2 * i + 1]);
    }
    // common scale keeps tiles square in space
    double span = std::max(maxx - minx, maxy - miny);
    double s = span > 0 ? 65535.0 / span : 0.0;
    std::vector<std::pair<uint64_t, int>> key(net.n);
    for (int i = 0; i < net.n; i++) {
        double fx = (net.pos[2 * i] - minx) * s, fy = (net.pos[2 * i + 1] - miny) * s;
        uint32_t qx = (uint32_t)std::min(65535.0, std::max(0.0, fx));
        uint32_t qy = (uint32_t)std::min(65535.0, std::max(0.0, fy));
        key[i] = {((uint64_t)(spread16(qx) | (spread16(qy) << 1)) << 32) | (uint32_t)i, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> order(net.n);
    for (int i = 0; i < net.n; i++) order[i] = key[i].second;
    return order;
}

struct Incidence {
    std::vector<int> brow, bent;  // bent = 2*bond + dir
    std::vector<int> trow, tent;  // triple index
};

Incidence incidence(const Network &net, bool bonds, bool triples)
{
    Incidence inc;
    inc.brow.assign(net.n + 1, 0);
    inc.trow.assign(net.n + 1, 0);
    if (bonds)
        for (int b = 0; b < net.e(); b++) {
            inc.brow[net.bonds[2 * b] + 1]++;
            inc.brow[net.bonds[2 * b + 1] + 1]++;
        }
    if (triples)
        for (int t = 0; t < net.t(); t++) {
            int h = net.tri[3 * t], a = net.tri[3 * t + 1], c = net.tri[3 * t + 2];
            inc.trow[h + 1]++;
            if (a != h) inc.trow[a + 1]++;
            if (c != h && c != a) inc.trow[c + 1]++;
        }
    for (int i = 0; i < net.n; i++) {
        inc.brow[i + 1] += inc.brow[i];
        inc.trow[i + 1] += inc.trow[i];
    }
    inc.bent.resize(inc.brow[net.n]);
    inc.tent.resize(inc.trow[net.n]);
    std::vector<int> fb(inc.brow.begin(), inc.brow.end() - 1);
    std::vector<int> ft(inc.trow.begin(), inc.trow.end() - 1);
    if (bonds)
        for (int b = 0; b < net.e(); b++) {
            inc.bent[fb[net.bonds[2 * b]]++] = 2 * b;
            inc.bent[fb[net.bonds[2 * b + 1]]++] = 2 * b + 1;
        }
    if (triples)
        for (int t = 0; t < net.t(); t++) {
            int h = net.tri[3 * t], a = net.tri[3 * t + 1], c = net.tri[3 * t + 2];
            inc.tent[ft[h]++] = t;
            if (a != h) inc.tent[ft[a]++] = t;
            if (c != h && c != a) inc.tent[ft[c]++] = t;
        }
    return inc;
}

inline bool wrap_ok(int w) { return w >= -kWrapBias && w < kWrapBias; }

}  // namespace

int build_plan(const std::vector<const Network *> &nets,
               const std::vector<double> &ks, const std::vector<double> &kb,
               int tile_nodes, int lds_budget, Plan &plan, std::string &err)
{
    plan = Plan();
    plan.tile_nodes = tile_nodes;
    plan.lds_budget = lds_budget;
    for (size_t k = 0; k < nets.size(); k++)
        if (kb[k] > 0 && !nets[k]->tkb.empty()) plan.has_tkb = true;

    int64_t node_base = 0;
    plan.brow.push_back(0);
    plan.trow.push_back(0);
    for (size_t k = 0; k < nets.size(); k++) {
        const Network &net = *nets[k];
        const bool use_b = ks[k] > 0, use_t = kb[k] > 0;
        if (use_t && !net.triples_built) {
            err = "bending stiffness > 0 but triples were not built";
            return NR_E_ARG;
        }
        PlanNet pn;
        pn.node_base = (int)node_base;
        pn.n_nodes = net.n;
        pn.n_bonds = use_b ? net.e() : 0;
        pn.n_triples = use_t ? net.t() : 0;
        pn.perm = spatial_order(net);
        pn.iperm.assign(net.n, 0);
        for (int i = 0; i < net.n; i++) pn.iperm[pn.perm[i]] = i;
        Incidence inc = incidence(net, use_b, use_t);
        pn.tile_base = (int)plan.tiles.size();

        // tiles: consecutive device nodes, split until the LDS budget holds
        std::vector<std::pair<int, int>> pending;  // [start, end) device-local
        for (int s = 0; s < net.n; s += tile_nodes)
            pending.push_back({s, std::min(net.n, s + tile_nodes)});
        std::vector<std::pair<int, int>> work(pending.rbegin(), pending.rend());
        std::vector<int> localmap(net.n, -1);
        while (!work.empty()) {
            auto [s, e] = work.back();
            work.pop_back();
            // own nodes in device order; halo = other nodes reached
            std::vector<int> halo;
            int nb = 0, nt = 0;
            for (int d = s; d < e; d++) localmap[pn.perm[d]] = d - s;
            auto touch = [&](int orig) {
                if (localmap[orig] == -1) {
                    localmap[orig] = -2;  // mark; real index assigned after sort
                    halo.push_back(pn.iperm[orig]);
                }
            };
            for (int d = s; d < e; d++) {
                int o = pn.perm[d];
                for (int q = inc.brow[o]; q < inc.brow[o + 1]; q++) {
                    int b = inc.bent[q] >> 1;
                    touch(net.bonds[2 * b]);
                    touch(net.bonds[2 * b + 1]);
                    nb++;
                }
                for (int q = inc.trow[o]; q < inc.trow[o + 1]; q++) {
                    int t = inc.tent[q];
                    touch(net.tri[3 * t]);
                    touch(net.tri[3 * t + 1]);
                    touch(net.tri[3 * t + 2]);
                    nt++;
                }
            }
            int local = (e - s) + (int)halo.size();
            int64_t need = Plan::tile_lds(local, nb, nt, plan.has_tkb);
            if ((need > lds_budget || local > 65535) && e - s > 1) {
                for (int d = s; d < e; d++) localmap[pn.perm[d]] = -1;
                for (int hd : halo) localmap[pn.perm[hd]] = -1;
                int mid = s + (e - s) / 2;
                work.push_back({mid, e});
                work.push_back({s, mid});
                continue;
            }
            if (need > lds_budget || local > 65535) {
                err = "a single node's incidence exceeds the LDS budget";
                for (int d = s; d < e; d++) localmap[pn.perm[d]] = -1;
                for (int hd : halo) localmap[pn.perm[hd]] = -1;
                return NR_E_TOPO;
            }
            std::sort(halo.begin(), halo.end());
            for (size_t q = 0; q < halo.size(); q++)
                localmap[pn.perm[halo[q]]] = (e - s) + (int)q;

            TileDev td;
            std::memset(&td, 0, sizeof(td));
            td.net = (int)k;
            td.node0 = (int)(node_base + s);
            td.n_own = e - s;
            td.n_halo = (int)halo.size();
            td.halo_off = (int)plan.halo.size();
            td.b0 = (int)plan.bcode.size();
            td.t0 = (int)plan.tcode.size();
            for (int hd : halo) plan.halo.push_back((int)(node_base + hd));
            for (int d = s; d < e; d++) {
                int o = pn.perm[d];
                for (int q = inc.brow[o]; q < inc.brow[o + 1]; q++) {
                    int b = inc.bent[q] >> 1, dir = inc.bent[q] & 1;
                    int other = net.bonds[2 * b + (dir ? 0 : 1)];
                    int wx = net.bwrap[2 * b], wy = net.bwrap[2 * b + 1];
                    if (!wrap_ok(wx) || !wrap_ok(wy)) {
                        err = "bond wrap code outside [-8, 7]";
                        return NR_E_TOPO;
                    }
                    uint32_t code = (uint32_t)localmap[other] |
                                    ((uint32_t)(wx + kWrapBias) << 16) |
                                    ((uint32_t)(wy + kWrapBias) << 20) | ((uint32_t)dir << 24);
                    plan.bcode.push_back(code);
                    plan.blen.push_back(net.rest[b]);
                }
                plan.brow.push_back((int)plan.bcode.size());
                for (int q = inc.trow[o]; q < inc.trow[o + 1]; q++) {
                    int t = inc.tent[q];
                    int w[4];
                    for (int z = 0; z < 4; z++) {
                        w[z] = net.twrap[4 * t + z];
                        if (!wrap_ok(w[z])) {
                            err = "triple wrap code outside [-8, 7]";
                            return NR_E_TOPO;
                        }
                    }
                    uint64_t code = (uint64_t)localmap[net.tri[3 * t]] |
                                    ((uint64_t)localmap[net.tri[3 * t + 1]] << 16) |
                                    ((uint64_t)localmap[net.tri[3 * t + 2]] << 32) |
                                    ((uint64_t)(w[0] + kWrapBias) << 48) |
                                    ((uint64_t)(w[1] + kWrapBias) << 52) |
                                    ((uint64_t)(w[2] + kWrapBias) << 56) |
                                    ((uint64_t)(w[3] + kWrapBias) << 60);
                    plan.tcode.push_back(code);
                    if (plan.has_tkb) plan.tkb.push_back(net.tkb.empty() ? kb[k] : net.tkb[t]);
                }
                plan.trow.push_back((int)plan.tcode.size());
            }
            td.nb = (int)plan.bcode.size() - td.b0;
            td.nt = (int)plan.tcode.size() - td.t0;
            plan.max_local = std::max(plan.max_local, local);
            plan.max_own = std::max(plan.max_own, td.n_own);
            plan.max_nb = std::max(plan.max_nb, td.nb);
            plan.max_nt = std::max(plan.max_nt, td.nt);
            plan.tiles.push_back(td);
            for (int d = s; d < e; d++) localmap[pn.perm[d]] = -1;
            for (int hd : halo) localmap[pn.perm[hd]] = -1;
        }
        pn.n_tiles = (int)plan.tiles.size() - pn.tile_base;
        for (int d = 0; d < net.n; d++) plan.mask.push_back(net.fcode[pn.perm[d]]);
        node_base += net.n;
        plan.nets.push_back(std::move(pn));
    }
    plan.n_total = node_base;
    if (plan.n_total > (int64_t)1 << 30 || plan.bcode.size() > (size_t)1 << 30 ||
        plan.tcode.size() > (size_t)1 << 30) {
        err = "batch too large for 32-bit device offsets";
        return NR_E_TOPO;
    }
    return NR_OK;
}

}  // namespace nr
