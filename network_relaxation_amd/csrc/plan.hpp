// Device layout of a batch of networks: spatial tiles with halos and
// per-node incidence lists (CSR) whose entries use tile-local node indices.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "network.hpp"

namespace nr {

// One tile: a run of consecutive device nodes of one network (its "own"
// nodes) plus the halo nodes its elements reach.  Local index space of a
// tile: [0, n_own) own nodes, [n_own, n_own + n_halo) halo nodes.
struct TileDev {
    int net;
    int node0;     // first own node, global device index
    int n_own;
    int n_halo;
    int halo_off;  // into the halo id array
    int b0, nb;    // bond entries [b0, b0 + nb)
    int t0, nt;    // triple entries [t0, t0 + nt)
    int pad[3];
};
static_assert(sizeof(TileDev) == 48, "TileDev layout");

// Bond entry code (u32): bits 0-15 neighbour (local), 16-19 wx+8, 20-23 wy+8,
// bit 24 = 1 when this node is the bond's second node.
// Triple entry code (u64): bits 0-15 hinge, 16-31 end a, 32-47 end b (local),
// 48-51 wx_a+8, 52-55 wy_a+8, 56-59 wx_b+8, 60-63 wy_b+8.
constexpr int kWrapBias = 8;

struct PlanNet {
    int node_base = 0, n_nodes = 0;
    int tile_base = 0, n_tiles = 0;
    int n_bonds = 0, n_triples = 0;  // elements on the device (0 when k == 0)
    std::vector<int> perm;   // device-local index -> original node
    std::vector<int> iperm;  // original node -> device-local index
};

struct Plan {
    int tile_nodes = 512;
    int lds_budget = 64 * 1024;
    bool has_tkb = false;
    int max_local = 0, max_nb = 0, max_nt = 0, max_own = 0;
    std::vector<PlanNet> nets;
    std::vector<TileDev> tiles;
    std::vector<int> brow, trow;           // N_total + 1, absolute entry offsets
    std::vector<uint32_t> bcode;
    std::vector<double> blen;
    std::vector<uint64_t> tcode;
    std::vector<double> tkb;               // per triple entry (has_tkb only)
    std::vector<int> halo;
    std::vector<unsigned char> mask;       // device order
    int64_t n_total = 0;

    // LDS bytes a tile needs in the FP64 kernel.
    static int64_t tile_lds(int local, int nb, int nt, bool tkb)
    {
        return 16ll * local + 12ll * nb + (tkb ? 16ll : 8ll) * nt;
    }
    int64_t lds_bytes() const { return tile_lds(max_local, max_nb, max_nt, has_tkb); }
};

// Builds the plan for nets[i] with stiffness ks[i], kb[i] (elements of a
// kind are dropped when its stiffness is 0, as the reference skips their
// kernels).  Returns an NR_ status.
int build_plan(const std::vector<const Network *> &nets,
               const std::vector<double> &ks, const std::vector<double> &kb,
               int tile_nodes, int lds_budget, Plan &plan, std::string &err);

}  // namespace nr
