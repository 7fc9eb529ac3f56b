// Device layout of a batch of networks: spatial tiles with halos.  Each tile's
// topology is one 16-byte aligned blob (copied verbatim into LDS by the
// kernel): halo node ids, bond and triple records with tile-local node
// indices, and the node -> element incidence lists (CSR) in element order.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "network.hpp"

namespace nr {

constexpr int kWrapBias = 8;

// The element phase of the kernel writes force vectors into one LDS array:
// [0, nb) the bond forces s r on their first node, then per triple the force
// on end a ([nb, nb+nt)), on end b ([nb+nt, nb+2nt)) and on the hinge
// ([nb+2nt, nb+3nt)).  An incidence entry (u16) is the index of the vector
// its node adds, with bit 15 set when the node adds its negation (the second
// node of a bond).
constexpr uint16_t kIncNegate = 0x8000;
constexpr int kMaxForceVectors = 0x7FFF;

// One tile: a run of consecutive device nodes of one network (its "own"
// nodes) plus the halo nodes its elements reach.  Local index space:
// [0, n_own) own nodes, [n_own, n_own + n_halo) halo nodes.
struct TileDev {
    int net;
    int node0;       // first own node, global device index
    int n_own, n_halo;
    int nb, nt;      // bond / triple records (elements touching own nodes)
    int n_inc;       // incidence entries
    int blob16;      // blob offset in 16-byte units
    int blob_bytes;  // multiple of 16
    // set by the relax kernel in its LDS copy of the descriptor (0 here):
    // where the tile's halo ids and its resident blob are held in LDS
    int lds_halo, lds_blob;
    int pad;
};
static_assert(sizeof(TileDev) == 48, "TileDev layout");

// Section offsets of a tile blob, in bytes (all 16-byte aligned):
//   halo   int32[n_halo]   global device node ids
//   bij    u32[nb]         i | j << 16            (local indices)
//   bw     u8[nb]          (wx + 8) | (wy + 8) << 4
//   bl     f64[nb]         rest length
//   tc     u64[nt]         h | a << 16 | b << 32 | wraps << 48 (4 nibbles,
//                          wx_a, wy_a, wx_b, wy_b, each + 8)
//   tk     f64[nt]         bending stiffness of the triple
//   ioff   u16[n_own + 1]  incidence list start per own node
//   inc    u16[n_inc]      force-vector index (| kIncNegate), element order
struct BlobLayout {
    int halo, bij, bw, bl, tc, tk, ioff, inc, bytes;
};

#ifdef __HIPCC__
#define NR_HD __host__ __device__
#else
#define NR_HD
#endif

NR_HD inline int align16(int x) { return (x + 15) & ~15; }

NR_HD inline BlobLayout blob_layout(int n_own, int n_halo, int nb, int nt, int n_inc)
{
    BlobLayout L;
    int o = 0;
    L.halo = o; o += (4 * n_halo + 15) & ~15;
    L.bij = o;  o += (4 * nb + 15) & ~15;
    L.bw = o;   o += (nb + 15) & ~15;
    L.bl = o;   o += (8 * nb + 15) & ~15;
    L.tc = o;   o += (8 * nt + 15) & ~15;
    L.tk = o;   o += (8 * nt + 15) & ~15;
    L.ioff = o; o += (2 * (n_own + 1) + 15) & ~15;
    L.inc = o;  o += (2 * n_inc + 15) & ~15;
    L.bytes = o;
    return L;
}

// LDS a tile needs beyond its blob: an FP64 pass holds positions (16 B per
// local node) and force vectors (16 B per bond, 48 B per triple); an
// FP32-stage pass FP64 anchors and FP32 displacements (24 B per local node)
// and FP32 force vectors (8 B per bond, 24 B per triple).
NR_HD inline int tile_work_bytes(int local, int nb, int nt)
{
    const int fv = nb + 3 * nt;
    const int w64 = 16 * local + 16 * fv, w32 = 24 * local + 8 * fv;
    return w64 > w32 ? w64 : w32;
}
inline int64_t tile_work_lds(int local, int nb, int nt) { return tile_work_bytes(local, nb, nt); }

struct PlanNet {
    int node_base = 0, n_nodes = 0;
    int tile_base = 0, n_tiles = 0;
    int n_bonds = 0, n_triples = 0;  // elements on the device (0 when k == 0)
    std::vector<int> perm;   // device-local index -> original node
    std::vector<int> iperm;  // original node -> device-local index
};

struct Plan {
    int tile_nodes = 512;
    int64_t lds_budget = 64 * 1024;
    int max_local = 0, max_nb = 0, max_nt = 0, max_own = 0, max_blob = 0, max_fv = 0;
    int64_t max_tile_lds = 0;  // largest blob + positions + force vectors of a tile
    std::vector<PlanNet> nets;
    std::vector<TileDev> tiles;
    std::vector<uint8_t> blob;             // all tile blobs, 16-byte aligned
    std::vector<unsigned char> mask;       // device order
    int64_t n_total = 0;

    int64_t lds_bytes() const { return max_tile_lds; }
};

// Builds the plan for nets[i] with stiffness ks[i], kb[i] (elements of a
// kind are dropped when its stiffness is 0, as the reference skips their
// kernels).  Returns an NR_ status.
int build_plan(const std::vector<const Network *> &nets,
               const std::vector<double> &ks, const std::vector<double> &kb,
               int tile_nodes, int64_t lds_budget, Plan &plan, std::string &err);

}  // namespace nr
