// Device layout builder: Morton-ordered spatial tiles, halos, element records
// with tile-local indices, and per-node incidence lists in element order
// (bonds by index, then triples by index), so that a node's force is summed
// in the same order as a sequential pass over the reference's bond kernel and
// then its triple kernel.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

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
    double span = std::max(maxx - minx, maxy - miny);
    double s = span > 0 ? 65535.0 / span : 0.0;
    std::vector<std::pair<uint64_t, int>> key(net.n);
    for (int i = 0; i < net.n; i++) {
        double fx = (net.pos[2 * i] - minx) * s, fy = (net.pos[2 * i + 1] - miny) * s;
        if (!(fx >= 0)) fx = 0;
        if (!(fy >= 0)) fy = 0;
        uint32_t qx = (uint32_t)std::min(65535.0, fx);
        uint32_t qy = (uint32_t)std::min(65535.0, fy);
        key[i] = {((uint64_t)(spread16(qx) | (spread16(qy) << 1)) << 32) | (uint32_t)i, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> order(net.n);
    for (int i = 0; i < net.n; i++) order[i] = key[i].second;
    return order;
}

// Per original node: incident bonds (2*bond + dir, bond order) and triples
// (triple order, once per distinct node).
struct Incidence {
    std::vector<int> brow, bent, trow, tent;
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

template <class T>
void put(std::vector<uint8_t> &blob, size_t base, int off, int idx, T v)
{
    std::memcpy(blob.data() + base + off + sizeof(T) * (size_t)idx, &v, sizeof(T));
}

}  // namespace

int build_plan(const std::vector<const Network *> &nets,
               const std::vector<double> &ks, const std::vector<double> &kb,
               int tile_nodes, int64_t lds_budget, Plan &plan, std::string &err)
{
    plan = Plan();
    plan.tile_nodes = tile_nodes;
    plan.lds_budget = lds_budget;
    int64_t node_base = 0;
    for (size_t k = 0; k < nets.size(); k++) {
        const Network &net = *nets[k];
        const bool use_b = ks[k] > 0, use_t = kb[k] > 0;
        if (net.n >= (1 << 27)) {
            // the kernel addresses a network's nodes with 32-bit byte offsets
            err = "a network has 2^27 nodes or more";
            return NR_E_TOPO;
        }
        if (use_t && !net.triples_built) {
            err = "bending stiffness > 0 but triples were not built";
            return NR_E_ARG;
        }
        for (int b = 0; use_b && b < net.e(); b++)
            if (!wrap_ok(net.bwrap[2 * b]) || !wrap_ok(net.bwrap[2 * b + 1])) {
                err = "bond wrap code outside [-8, 7]";
                return NR_E_TOPO;
            }
        for (int t = 0; use_t && t < net.t(); t++)
            for (int z = 0; z < 4; z++)
                if (!wrap_ok(net.twrap[4 * t + z])) {
                    err = "triple wrap code outside [-8, 7]";
                    return NR_E_TOPO;
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

        // [start, end) device-local chunks as a stack, the first chunk on top
        std::vector<std::pair<int, int>> work;
        const int n_chunks = (net.n + tile_nodes - 1) / tile_nodes;
        for (int q = n_chunks - 1; q >= 0; q--)
            work.push_back({q * tile_nodes, std::min(net.n, (q + 1) * tile_nodes)});
        std::vector<int> localmap(net.n, -1), bslot(use_b ? net.e() : 0, -1),
            tslot(use_t ? net.t() : 0, -1);
        while (!work.empty()) {
            auto [s, e] = work.back();
            work.pop_back();
            for (int d = s; d < e; d++) localmap[pn.perm[d]] = d - s;
            std::vector<int> halo, bl, tl;
            int n_inc = 0;
            auto touch = [&](int orig) {
                if (localmap[orig] == -1) {
                    localmap[orig] = -2;
                    halo.push_back(pn.iperm[orig]);
                }
            };
            for (int d = s; d < e; d++) {
                int o = pn.perm[d];
                for (int q = inc.brow[o]; q < inc.brow[o + 1]; q++) {
                    int b = inc.bent[q] >> 1;
                    n_inc++;
                    if (bslot[b] == -1) {
                        bslot[b] = -2;
                        bl.push_back(b);
                        touch(net.bonds[2 * b]);
                        touch(net.bonds[2 * b + 1]);
                    }
                }
                for (int q = inc.trow[o]; q < inc.trow[o + 1]; q++) {
                    int t = inc.tent[q];
                    int h = net.tri[3 * t], a = net.tri[3 * t + 1], c = net.tri[3 * t + 2];
                    n_inc += (o == h) + (o == a) + (o == c);
                    if (tslot[t] == -1) {
                        tslot[t] = -2;
                        tl.push_back(t);
                        touch(h);
                        touch(a);
                        touch(c);
                    }
                }
            }
            const int n_own = e - s, n_halo = (int)halo.size(), nb = (int)bl.size(),
                      nt = (int)tl.size(), local = n_own + n_halo;
            BlobLayout Lb = blob_layout(n_own, n_halo, nb, nt, n_inc);
            int64_t need = Lb.bytes + tile_work_lds(local, nb, nt);
            bool fits = need <= lds_budget && local <= 65535 &&
                        nb + 3 * nt <= kMaxForceVectors && n_inc <= 65535;
            auto clear = [&]() {
                for (int d = s; d < e; d++) localmap[pn.perm[d]] = -1;
                for (int hd : halo) localmap[pn.perm[hd]] = -1;
                for (int b : bl) bslot[b] = -1;
                for (int t : tl) tslot[t] = -1;
            };
            if (!fits) {
                clear();
                if (e - s <= 1) {
                    err = "a single node's incidence exceeds the LDS budget";
                    return NR_E_TOPO;
                }
                int mid = s + (e - s) / 2;
                work.push_back({mid, e});
                work.push_back({s, mid});
                continue;
            }
            std::sort(halo.begin(), halo.end());
            for (size_t q = 0; q < halo.size(); q++) localmap[pn.perm[halo[q]]] = n_own + (int)q;
            std::sort(bl.begin(), bl.end());
            std::sort(tl.begin(), tl.end());
            for (size_t q = 0; q < bl.size(); q++) bslot[bl[q]] = (int)q;
            for (size_t q = 0; q < tl.size(); q++) tslot[tl[q]] = (int)q;

            TileDev td;
            std::memset(&td, 0, sizeof(td));
            td.net = (int)k;
            td.node0 = (int)(node_base + s);
            td.n_own = n_own;
            td.n_halo = n_halo;
            td.nb = nb;
            td.nt = nt;
            td.n_inc = n_inc;
            td.blob_bytes = Lb.bytes;
            size_t base = plan.blob.size();
            td.blob16 = (int)(base / 16);
            plan.blob.resize(base + Lb.bytes, 0);
            for (int q = 0; q < n_halo; q++)
                put<int32_t>(plan.blob, base, Lb.halo, q, (int32_t)(node_base + halo[q]));
            for (int q = 0; q < nb; q++) {
                int b = bl[q];
                uint32_t ij = (uint32_t)localmap[net.bonds[2 * b]] |
                              ((uint32_t)localmap[net.bonds[2 * b + 1]] << 16);
                uint8_t w = (uint8_t)((net.bwrap[2 * b] + kWrapBias) |
                                      ((net.bwrap[2 * b + 1] + kWrapBias) << 4));
                put<uint32_t>(plan.blob, base, Lb.bij, q, ij);
                put<uint8_t>(plan.blob, base, Lb.bw, q, w);
                put<double>(plan.blob, base, Lb.bl, q, net.rest[b]);
            }
            for (int q = 0; q < nt; q++) {
                int t = tl[q];
                uint64_t code = (uint64_t)localmap[net.tri[3 * t]] |
                                ((uint64_t)localmap[net.tri[3 * t + 1]] << 16) |
                                ((uint64_t)localmap[net.tri[3 * t + 2]] << 32) |
                                ((uint64_t)(net.twrap[4 * t] + kWrapBias) << 48) |
                                ((uint64_t)(net.twrap[4 * t + 1] + kWrapBias) << 52) |
                                ((uint64_t)(net.twrap[4 * t + 2] + kWrapBias) << 56) |
                                ((uint64_t)(net.twrap[4 * t + 3] + kWrapBias) << 60);
                put<uint64_t>(plan.blob, base, Lb.tc, q, code);
                put<double>(plan.blob, base, Lb.tk, q, net.tkb.empty() ? kb[k] : net.tkb[t]);
            }
            int ii = 0;
            for (int d = s; d < e; d++) {
                int o = pn.perm[d];
                put<uint16_t>(plan.blob, base, Lb.ioff, d - s, (uint16_t)ii);
                for (int q = inc.brow[o]; q < inc.brow[o + 1]; q++) {
                    int b = inc.bent[q] >> 1, dir = inc.bent[q] & 1;
                    put<uint16_t>(plan.blob, base, Lb.inc, ii++,
                                  (uint16_t)(bslot[b] | (dir ? kIncNegate : 0)));
                }
                // a node that is several corners of one triple adds hinge, end a,
                // end b in that order (the reference's atomics at :461-468)
                for (int q = inc.trow[o]; q < inc.trow[o + 1]; q++) {
                    int t = inc.tent[q];
                    int h = net.tri[3 * t], a = net.tri[3 * t + 1], c = net.tri[3 * t + 2];
                    const int ts = tslot[t];
                    if (o == h) put<uint16_t>(plan.blob, base, Lb.inc, ii++, (uint16_t)(nb + 2 * nt + ts));
                    if (o == a) put<uint16_t>(plan.blob, base, Lb.inc, ii++, (uint16_t)(nb + ts));
                    if (o == c) put<uint16_t>(plan.blob, base, Lb.inc, ii++, (uint16_t)(nb + nt + ts));
                }
            }
            put<uint16_t>(plan.blob, base, Lb.ioff, n_own, (uint16_t)ii);
            plan.max_local = std::max(plan.max_local, local);
            plan.max_own = std::max(plan.max_own, n_own);
            plan.max_nb = std::max(plan.max_nb, nb);
            plan.max_nt = std::max(plan.max_nt, nt);
            plan.max_blob = std::max(plan.max_blob, Lb.bytes);
            plan.max_fv = std::max(plan.max_fv, nb + 3 * nt);
            plan.max_tile_lds = std::max(plan.max_tile_lds, need);
            plan.tiles.push_back(td);
            clear();
        }
        pn.n_tiles = (int)plan.tiles.size() - pn.tile_base;
        for (int d = 0; d < net.n; d++) plan.mask.push_back(net.fcode[pn.perm[d]]);
        node_base += net.n;
        plan.nets.push_back(std::move(pn));
    }
    plan.n_total = node_base;
    if (plan.n_total > ((int64_t)1 << 30) || plan.blob.size() / 16 > ((size_t)1 << 31) - 1) {
        err = "batch too large for 32-bit device offsets";
        return NR_E_TOPO;
    }
    return NR_OK;
}

}  // namespace nr
