// Network + FEM-gel composite on the device: the C-ABI of nr_mesh /
// nr_composite and the batch protocol of the reference's composite program
// (src/network_mesh_relax.cu: relaxation_batch :1393-1421, do_relaxation_run
// :1095-1391, run_fire_md :865-1066).
//
// A relax uploads the FIRE scalars, zeroes v, computes the first force and
// drift, then replays a captured chunk of kMeshChunk steps (two kernels
// each, mesh_kernels.hip: the fused element + gather pass over spatial tiles
// and the FIRE update) until the device's done flag is up; positions stay
// resident between the increments of a ramp.  Points live on the device in
// Morton order (positions and forces are permuted at the C ABI).  With
// NRELAX_GRAPHS=0 in the environment at nr_composite_create, each chunk's
// kernels are launched directly instead (same kernels, same order; for tools
// that do not follow graph launches); with NRELAX_MESH_TILES=0 a step runs
// the unfused element and gather kernels (three launches, the same forces
// bit for bit).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nrelax.h"
#include "internal.hpp"
#include "mesh.hpp"
#include "mesh_device.hpp"
#include "protocol.hpp"

struct nr_mesh {
    nr::Mesh m;
};

namespace {

constexpr int kMeshChunk = 32;  // FIRE steps per captured graph (even)
// Dynamic LDS of a composite tile: two workgroups of mesh_tile share a CU's
// 160 KB (with their static LDS).
constexpr int kMeshTileLds = 76 * 1024;

}  // namespace

struct nr_composite {
    nr::Mesh mesh;
    nr::MeshSetup setup;
    nr_gel_params gel{};
    int device = 0;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    nr::DevBuf<double> x, v, f, part, tot, state, epart, inv, area;
    nr::DevBuf<unsigned> ticket;
    nr::DevBuf<double2> slot, bond_geo, tri_off;
    nr::DevBuf<int2> bond;
    nr::DevBuf<int4> tri, facet;
    nr::DevBuf<int> ent, cnt, twin;
    nr::DevBuf<unsigned char> fcode;
    std::vector<int> dev_of;  // device index of each point of the file
    // fused element + gather kernel
    nr::DevBuf<nr::MeshTile> tiles;
    nr::DevBuf<uint4> tblob;
    nr::DevBuf<unsigned> tb_ij;
    nr::DevBuf<double2> tb_geo, tt_off, tf_m;
    nr::DevBuf<uint2> tt_idx, tf_idx;
    nr::DevBuf<double> tf_area;
    nr::MeshDev md{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    nr_relax_params captured{};
    bool have_graph = false;
    bool graphs = true;  // false: each chunk's kernels launched directly
    int energy_blocks = 1;
    double last_ms = 0.0;

    ~nr_composite()
    {
        // a relax that returned early (an error, the step bound) may have
        // left graph replays queued: they finish before the graph goes
        if (st) (void)hipStreamSynchronize(st);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace nr {
namespace {

// Slot entries per point, (slot << 1) | negate, in element order, in three
// sections: bonds (bond b is slot b, + on its first point, - on its second),
// triples (triple t is slots nb + t for end a and nb + nt + t for end b,
// both negated on the hinge) and facets (facet f is slot
// nb + 2nt + k nf + f on its corner k).  Each section is column-major with the section's largest
// per-point count as its width: entry k of point i at sec[s] + k * n + i.
// cnt packs the three counts of a point in 8 bits each.
int build_gather(int n, const std::vector<int> &bonds, int nb, const std::vector<int> &tri, int nt,
                 const std::vector<int> &facets, int nf, std::vector<int> &ent,
                 std::vector<int> &cnt, long sec[3])
{
    std::vector<std::vector<int>> lists(3 * (size_t)n);
    auto put = [&](int s, int point, int e) { lists[3 * (size_t)point + s].push_back(e); };
    for (int b = 0; b < nb; b++) {
        put(0, bonds[2 * b], b << 1);
        put(0, bonds[2 * b + 1], (b << 1) | 1);
    }
    for (int t = 0; t < nt; t++) {
        const int sa = nb + t, sb = nb + nt + t;
        put(1, tri[3 * t], (sa << 1) | 1);
        put(1, tri[3 * t], (sb << 1) | 1);
        put(1, tri[3 * t + 1], sa << 1);
        put(1, tri[3 * t + 2], sb << 1);
    }
    for (int f = 0; f < nf; f++)
        for (int k = 0; k < 3; k++) put(2, facets[3 * f + k], (nb + 2 * nt + k * nf + f) << 1);
    size_t width[3] = {0, 0, 0};
    for (int i = 0; i < n; i++)
        for (int s = 0; s < 3; s++)
            width[s] = std::max(width[s], lists[3 * (size_t)i + s].size());
    if (width[0] > 255 || width[1] > 255 || width[2] > 255)
        return fail(NR_E_TOPO, "a point has more than 255 entries of one element type");
    sec[0] = 0;
    sec[1] = (long)(width[0] * (size_t)n);
    sec[2] = sec[1] + (long)(width[1] * (size_t)n);
    ent.assign((size_t)sec[2] + width[2] * (size_t)n + 1, 0);
    cnt.assign((size_t)n, 0);
    for (int i = 0; i < n; i++) {
        for (int s = 0; s < 3; s++) {
            const std::vector<int> &l = lists[3 * (size_t)i + s];
            for (size_t k = 0; k < l.size(); k++) ent[(size_t)sec[s] + k * (size_t)n + i] = l[k];
        }
        cnt[i] = (int)lists[3 * (size_t)i].size() | (int)lists[3 * (size_t)i + 1].size() << 8 |
                 (int)lists[3 * (size_t)i + 2].size() << 16;
    }
    return NR_OK;
}

// Host side of the fused element + gather kernel: tiles of MeshTile with
// their blobs and tile-major element records.
struct TileArrays {
    std::vector<MeshTile> tiles;
    std::vector<uint8_t> blob;
    std::vector<unsigned> b_ij;
    std::vector<double2> b_geo;
    std::vector<uint2> t_idx;
    std::vector<double2> t_off;
    std::vector<uint2> f_idx;
    std::vector<double2> f_m;
    std::vector<double> f_area;
    int max_lds = 0;
};

uint32_t spread_bits(uint32_t x)
{
    x &= 0xFFFF;
    x = (x | (x << 8)) & 0x00FF00FF;
    x = (x | (x << 4)) & 0x0F0F0F0F;
    x = (x | (x << 2)) & 0x33333333;
    x = (x | (x << 1)) & 0x55555555;
    return x;
}

// Device order of the composite's points: Morton order of their positions
// (16 bits per axis, ties by index), so that a tile's points are one
// contiguous run of the state arrays and its neighbours lie close by.
std::vector<int> morton_order(int n, const std::vector<double> &pos)
{
    std::vector<int> order((size_t)n);
    if (n <= 0) return order;
    double x0 = pos[0], x1 = pos[0], y0 = pos[1], y1 = pos[1];
    for (int i = 0; i < n; i++) {
        x0 = std::min(x0, pos[2 * i]);
        x1 = std::max(x1, pos[2 * i]);
        y0 = std::min(y0, pos[2 * i + 1]);
        y1 = std::max(y1, pos[2 * i + 1]);
    }
    const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
    std::vector<std::pair<uint32_t, int>> key((size_t)n);
    for (int i = 0; i < n; i++) {
        const uint32_t qx = (uint32_t)((pos[2 * i] - x0) * sx);
        const uint32_t qy = (uint32_t)((pos[2 * i + 1] - y0) * sy);
        key[i] = {spread_bits(qx) | spread_bits(qy) << 1, i};
    }
    std::sort(key.begin(), key.end());
    for (int i = 0; i < n; i++) order[i] = key[i].second;
    return order;
}

// Points (in device order) cut into runs of tile_points (halved until a
// tile's LDS fits lds_budget and its elements the registers of mesh_tile).
// A tile takes
// every element incident to an own point or to its twin (twin non-empty:
// coupled twins), in global element order per type; its points are the own
// ones, then the others those elements reach in order of first use; each own
// point's entries follow build_gather's order (bonds, triples, facets, each
// in element order), then its twin's.
int build_tiles(int n, const std::vector<int> &bonds,
                const std::vector<double2> &geo, int nb, const std::vector<int> &tri,
                const std::vector<double2> &toff, int nt, const std::vector<int> &facets,
                const std::vector<double> &inv, const std::vector<double> &area, int nf,
                const std::vector<int> &twin, const std::vector<unsigned char> &fcode,
                int lds_budget, int tile_points, TileArrays &T)
{
    T = TileArrays();
    if (n <= 0) return NR_OK;
    // incident elements per point, in element order: bond 2b | second,
    // triple 3t + role (0 hinge, 1 end a, 2 end b), facet 3f + corner
    std::vector<std::vector<int>> binc((size_t)n), tinc((size_t)n), finc((size_t)n);
    for (int b = 0; b < nb; b++) {
        binc[bonds[2 * b]].push_back(2 * b);
        binc[bonds[2 * b + 1]].push_back(2 * b + 1);
    }
    for (int t = 0; t < nt; t++)
        for (int r = 0; r < 3; r++) tinc[tri[3 * t + r]].push_back(3 * t + r);
    for (int f = 0; f < nf; f++)
        for (int k = 0; k < 3; k++) finc[facets[3 * f + k]].push_back(3 * f + k);
    std::vector<int> lstamp((size_t)n, -1), loc((size_t)n, 0);
    std::vector<int> bstamp((size_t)nb, -1), tstamp((size_t)nt, -1), fstamp((size_t)nf, -1);
    std::vector<int> bl((size_t)nb, 0), tl((size_t)nt, 0), fl((size_t)nf, 0);
    int stamp = 0;
    // ranges of the Morton order, in order; one that does not fit is halved
    // in place
    std::vector<std::pair<int, int>> ranges;
    for (int a = 0; a < n; a += tile_points) ranges.push_back({a, std::min(n, a + tile_points)});
    std::vector<int> own, halo, blist, tlist, flist, ent;
    std::vector<uint16_t> ioff, imid;
    std::vector<uint8_t> mask;
    for (size_t ri = 0; ri < ranges.size(); ri++) {
        const int a = ranges[ri].first, e = ranges[ri].second;
        const int tag = stamp++;
        own.clear();
        halo.clear();
        blist.clear();
        tlist.clear();
        flist.clear();
        for (int k = a; k < e; k++) {
            const int p = k;
            own.push_back(p);
            lstamp[p] = tag;
            loc[p] = k - a;
        }
        auto collect = [&](int q) {
            for (int c : binc[q])
                if (bstamp[c >> 1] != tag) {
                    bstamp[c >> 1] = tag;
                    blist.push_back(c >> 1);
                }
            for (int c : tinc[q])
                if (tstamp[c / 3] != tag) {
                    tstamp[c / 3] = tag;
                    tlist.push_back(c / 3);
                }
            for (int c : finc[q])
                if (fstamp[c / 3] != tag) {
                    fstamp[c / 3] = tag;
                    flist.push_back(c / 3);
                }
        };
        for (int p : own) {
            collect(p);
            if (!twin.empty() && twin[p] >= 0) collect(twin[p]);
        }
        std::sort(blist.begin(), blist.end());
        std::sort(tlist.begin(), tlist.end());
        std::sort(flist.begin(), flist.end());
        auto reach = [&](int q) {
            if (lstamp[q] != tag) {
                lstamp[q] = tag;
                loc[q] = (int)(own.size() + halo.size());
                halo.push_back(q);
            }
        };
        for (size_t i = 0; i < blist.size(); i++) {
            bl[blist[i]] = (int)i;
            reach(bonds[2 * blist[i]]);
            reach(bonds[2 * blist[i] + 1]);
        }
        for (size_t i = 0; i < tlist.size(); i++) {
            tl[tlist[i]] = (int)i;
            for (int r = 0; r < 3; r++) reach(tri[3 * tlist[i] + r]);
        }
        for (size_t i = 0; i < flist.size(); i++) {
            fl[flist[i]] = (int)i;
            for (int k = 0; k < 3; k++) reach(facets[3 * flist[i] + k]);
        }
        const int n_own = (int)own.size(), n_loc = n_own + (int)halo.size();
        const int tb = (int)blist.size(), tt = (int)tlist.size(), tf = (int)flist.size();
        const long n_fv = (long)tb + 2L * tt + 3L * tf;
        auto entries = [&](int q) {
            for (int c : binc[q]) ent.push_back(bl[c >> 1] << 1 | (c & 1));
            for (int c : tinc[q]) {
                const int t = tl[c / 3], role = c % 3;
                if (role == 0) {
                    ent.push_back((tb + t) << 1 | 1);
                    ent.push_back((tb + tt + t) << 1 | 1);
                } else {
                    ent.push_back((role == 1 ? tb + t : tb + tt + t) << 1);
                }
            }
            for (int c : finc[q]) ent.push_back((tb + 2 * tt + (c % 3) * tf + fl[c / 3]) << 1);
        };
        ent.clear();
        ioff.assign((size_t)n_own + 1, 0);
        imid.assign((size_t)n_own, 0);
        mask.assign((size_t)n_own, 0);
        for (int k = 0; k < n_own; k++) {
            const int p = own[k];
            ioff[k] = (uint16_t)std::min<size_t>(ent.size(), 0xFFFF);
            entries(p);
            imid[k] = (uint16_t)std::min<size_t>(ent.size(), 0xFFFF);
            mask[k] = (uint8_t)(fcode[p] & 3);
            if (!twin.empty() && twin[p] >= 0) {
                entries(twin[p]);
                mask[k] |= (uint8_t)((fcode[twin[p]] & 3) << 2 | 16);
            }
        }
        ioff[n_own] = (uint16_t)std::min<size_t>(ent.size(), 0xFFFF);
        const MeshBlob B = mesh_blob(n_own, n_loc - n_own, (int)ent.size());
        const long lds = (long)B.bytes + 16L * n_loc + 16L * n_fv;
        const bool fits = lds <= lds_budget && n_loc <= kTileLocalPerThread * 256 && n_fv <= 32767 &&
                          ent.size() <= 65535 && B.bytes <= kTileBlobRows * 256 * 16 &&
                          tb <= kTileBondsPerThread * 256 &&
                          tt <= kTileTriplesPerThread * 256 && tf <= kTileFacetsPerThread * 256;
        if (!fits) {
            if (e - a <= 8) return fail(NR_E_TOPO, "a composite tile exceeds the LDS budget");
            // halve this range and redo it (the halves are processed next)
            const int mid = a + (e - a) / 2;
            ranges.insert(ranges.begin() + (long)ri + 1, {{a, mid}, {mid, e}});
            continue;
        }
        MeshTile t{};
        t.blob16 = (int)(T.blob.size() / 16);
        t.blob_bytes = B.bytes;
        t.n_own = n_own;
        t.n_halo = n_loc - n_own;
        t.n_inc = (int)ent.size();
        t.nb = tb;
        t.nt = tt;
        t.nf = tf;
        t.b0 = (int)T.b_ij.size();
        t.t0 = (int)T.t_idx.size();
        t.f0 = (int)T.f_idx.size();
        T.tiles.push_back(t);
        T.max_lds = std::max<int>(T.max_lds, (int)lds);
        const size_t base = T.blob.size();
        T.blob.resize(base + (size_t)B.bytes, 0);
        uint8_t *bp = T.blob.data() + base;
        int32_t *ids = reinterpret_cast<int32_t *>(bp + B.ids);
        for (int k = 0; k < n_own; k++) ids[k] = own[k];
        for (size_t k = 0; k < halo.size(); k++) ids[n_own + k] = halo[k];
        std::memcpy(bp + B.ioff, ioff.data(), 2 * ioff.size());
        std::memcpy(bp + B.imid, imid.data(), 2 * imid.size());
        std::memcpy(bp + B.mask, mask.data(), mask.size());
        uint16_t *ip = reinterpret_cast<uint16_t *>(bp + B.inc);
        for (size_t k = 0; k < ent.size(); k++) ip[k] = (uint16_t)ent[k];
        for (int b : blist) {
            T.b_ij.push_back((unsigned)loc[bonds[2 * b]] | (unsigned)loc[bonds[2 * b + 1]] << 16);
            T.b_geo.push_back(geo[b]);
        }
        for (int q : tlist) {
            T.t_idx.push_back(make_uint2((unsigned)loc[tri[3 * q]] | (unsigned)loc[tri[3 * q + 1]] << 16,
                                         (unsigned)loc[tri[3 * q + 2]]));
            T.t_off.push_back(toff[q]);
        }
        for (int f : flist) {
            T.f_idx.push_back(make_uint2(
                (unsigned)loc[facets[3 * f]] | (unsigned)loc[facets[3 * f + 1]] << 16,
                (unsigned)loc[facets[3 * f + 2]]));
            for (int k = 0; k < 3; k++)
                T.f_m.push_back(make_double2(inv[6 * (size_t)f + 2 * k], inv[6 * (size_t)f + 2 * k + 1]));
            T.f_area.push_back(area[f]);
        }
    }
    if (T.tiles.size() > (size_t)(1 << 30)) return fail(NR_E_TOPO, "too many composite tiles");
    return NR_OK;
}

int composite_upload(nr_composite &c)
{
    const Network &net = c.mesh.net;
    const MeshSetup &s = c.setup;
    const int n = net.n;
    // the reference skips a term whose modulus is 0 (:928-951)
    const int nb = c.gel.ks > 0 ? net.e() : 0;
    const int nt = c.gel.kb > 0 ? net.t() : 0;
    const int nf = c.gel.bmod > 0 ? c.mesh.nf() : 0;
    // device order of the points (Morton), and every point index in it;
    // elements keep the file's order
    const std::vector<int> order = morton_order(n, s.pos);
    c.dev_of.assign((size_t)n, 0);
    for (int d = 0; d < n; d++) c.dev_of[order[d]] = d;
    std::vector<int> dbonds(2 * (size_t)nb), dtri(3 * (size_t)nt), dfac(3 * (size_t)nf);
    for (size_t k = 0; k < dbonds.size(); k++) dbonds[k] = c.dev_of[net.bonds[k]];
    for (size_t k = 0; k < dtri.size(); k++) dtri[k] = c.dev_of[net.tri[k]];
    for (size_t k = 0; k < dfac.size(); k++) dfac[k] = c.dev_of[c.mesh.facets[k]];
    std::vector<double> dpos(2 * (size_t)n);
    std::vector<unsigned char> dfcode((size_t)n);
    for (int d = 0; d < n; d++) {
        dpos[2 * (size_t)d] = s.pos[2 * (size_t)order[d]];
        dpos[2 * (size_t)d + 1] = s.pos[2 * (size_t)order[d] + 1];
        dfcode[d] = net.fcode[order[d]];
    }
    std::vector<int> ent, cnt;
    long sec[3];
    int rc = build_gather(n, dbonds, nb, dtri, nt, dfac, nf, ent, cnt, sec);
    if (rc) return rc;
    std::vector<int2> bond((size_t)nb);
    std::vector<double2> geo((size_t)nb);
    for (int b = 0; b < nb; b++) {
        bond[b] = make_int2(dbonds[2 * b], dbonds[2 * b + 1]);
        geo[b] = make_double2(net.rest[b], s.width * net.bwrap[2 * b]);
    }
    std::vector<int4> tri((size_t)nt);
    std::vector<double2> toff((size_t)nt);
    for (int t = 0; t < nt; t++) {
        tri[t] = make_int4(dtri[3 * t], dtri[3 * t + 1], dtri[3 * t + 2], 0);
        toff[t] = make_double2(s.width * net.twrap[4 * t], s.width * net.twrap[4 * t + 2]);
    }
    std::vector<int4> fac((size_t)nf);
    for (int f = 0; f < nf; f++) fac[f] = make_int4(dfac[3 * f], dfac[3 * f + 1], dfac[3 * f + 2], 0);
    std::vector<double> inv(s.inv.begin(), s.inv.begin() + 6 * (size_t)nf);
    std::vector<double> area(c.mesh.areas.begin(), c.mesh.areas.begin() + nf);
    std::vector<int> twin;
    if (nf > 0 && c.mesh.np() > 0) {  // duplicate_forces runs with the gel only (:946-951)
        twin.assign((size_t)n, -1);
        for (int p = 0; p < c.mesh.np(); p++) {
            const int a = c.dev_of[c.mesh.twins[2 * p]], b = c.dev_of[c.mesh.twins[2 * p + 1]];
            twin[(size_t)a] = b;
            twin[(size_t)b] = a;
        }
    }
    const size_t nslot = (size_t)nb + 2 * (size_t)nt + 3 * (size_t)nf;
    if (nslot >= (1u << 30)) return fail(NR_E_TOPO, "composite too large for the slot index");
    HIP_TRY(c.x.upload(dpos));
    HIP_TRY(c.v.alloc(2 * (size_t)n));
    HIP_TRY(c.f.alloc(2 * (size_t)n));
    HIP_TRY(c.bond.upload(bond));
    HIP_TRY(c.bond_geo.upload(geo));
    HIP_TRY(c.tri.upload(tri));
    HIP_TRY(c.tri_off.upload(toff));
    HIP_TRY(c.facet.upload(fac));
    HIP_TRY(c.inv.upload(inv));
    HIP_TRY(c.area.upload(area));
    HIP_TRY(c.slot.alloc(nslot));
    HIP_TRY(c.ent.upload(ent));
    HIP_TRY(c.cnt.upload(cnt));
    if (!twin.empty()) HIP_TRY(c.twin.upload(twin));
    HIP_TRY(c.fcode.upload(dfcode));
    // the fused element + gather kernel (NRELAX_MESH_TILES=0 at create: the
    // element and gather kernels instead)
    TileArrays T;
    long tile_slots = 512;
    const char *t_env = std::getenv("NRELAX_MESH_TILES");
    if (!(t_env && std::strcmp(t_env, "0") == 0)) {
        // tiles of at most kMeshTilePoints points, as few rounds of the
        // resident workgroups (two per CU) as that allows, and the points
        // spread evenly over the rounds (no round of a few tiles at the end)
        int n_cu = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c.device) == hipSuccess && prop.multiProcessorCount > 0)
            n_cu = prop.multiProcessorCount;
        const long slots = 2L * n_cu;
        tile_slots = slots;
        const long rounds = std::max(1L, ((long)n + slots * kMeshTilePoints - 1) /
                                             (slots * kMeshTilePoints));
        const int tile_points = (int)std::min<long>(kMeshTilePoints,
                                                    std::max(32L, ((long)n + slots * rounds - 1) /
                                                                      (slots * rounds)));
        rc = build_tiles(n, dbonds, geo, nb, dtri, toff, nt, dfac, inv, area, nf, twin, dfcode,
                         kMeshTileLds, tile_points, T);
        if (rc) return rc;
        std::vector<uint4> blob16(T.blob.size() / 16);
        if (!T.blob.empty()) std::memcpy(blob16.data(), T.blob.data(), T.blob.size());
        HIP_TRY(c.tiles.upload(T.tiles));
        HIP_TRY(c.tblob.upload(blob16));
        HIP_TRY(c.tb_ij.upload(T.b_ij));
        HIP_TRY(c.tb_geo.upload(T.b_geo));
        HIP_TRY(c.tt_idx.upload(T.t_idx));
        HIP_TRY(c.tt_off.upload(T.t_off));
        HIP_TRY(c.tf_idx.upload(T.f_idx));
        HIP_TRY(c.tf_m.upload(T.f_m));
        HIP_TRY(c.tf_area.upload(T.f_area));
    }
    HIP_TRY(c.part.alloc(4 * std::max<size_t>(kMeshMaxBlocks, T.tiles.size())));
    HIP_TRY(c.tot.alloc(4));
    HIP_TRY(c.ticket.alloc(1));
    HIP_TRY(hipMemset(c.ticket.p, 0, sizeof(unsigned)));
    HIP_TRY(c.state.alloc(kStWords));
    HIP_TRY(c.epart.alloc(3 * (size_t)kMeshMaxBlocks));

    MeshDev &m = c.md;
    m = MeshDev{};
    m.n = n;
    m.nb = nb;
    m.nt = nt;
    m.nf = nf;
    m.x = c.x.p;
    m.v = c.v.p;
    m.f = c.f.p;
    m.bond = c.bond.p;
    m.bond_geo = c.bond_geo.p;
    m.tri = c.tri.p;
    m.tri_off = c.tri_off.p;
    m.facet = c.facet.p;
    m.inv = c.inv.p;
    m.area = c.area.p;
    m.slot = c.slot.p;
    m.ent = c.ent.p;
    m.cnt = c.cnt.p;
    for (int k = 0; k < 3; k++) m.sec[k] = sec[k];
    m.twin = twin.empty() ? nullptr : c.twin.p;
    m.fcode = c.fcode.p;
    m.part = c.part.p;
    m.tot = c.tot.p;
    m.ticket = c.ticket.p;
    m.st = c.state.p;
    m.gather_blocks = mesh_point_blocks(n);
    m.n_tiles = (int)T.tiles.size();
    m.tile_grid = (int)std::min<long>(m.n_tiles, tile_slots);
    m.tile_lds = T.max_lds;
    m.tiles = c.tiles.p;
    m.tblob = c.tblob.p;
    m.tb_ij = c.tb_ij.p;
    m.tb_geo = c.tb_geo.p;
    m.tt_idx = c.tt_idx.p;
    m.tt_off = c.tt_off.p;
    m.tf_idx = c.tf_idx.p;
    m.tf_m = c.tf_m.p;
    m.tf_area = c.tf_area.p;
    m.ks = c.gel.ks;
    m.kb = c.gel.kb;
    m.smod = c.gel.smod;
    m.bmod = c.gel.bmod;
    m.pref_d = s.pref_d;
    m.sqrt_dof = std::sqrt((double)(2 * n));
    c.energy_blocks = mesh_element_blocks(m);
    return NR_OK;
}

void set_fire(MeshDev &m, const nr_relax_params &p)
{
    m.alpha0 = p.alpha0;
    m.f_alpha = p.f_alpha;
    m.dt_max = p.dt_max;
    m.f_inc = p.f_inc;
    m.f_dec = p.f_dec;
    m.mass = p.mass;
    m.fcut = p.fcut;
    m.nmin = p.nmin;
    m.max_steps = p.max_steps;
    m.criterion = p.criterion;
}

bool same_params(const nr_relax_params &a, const nr_relax_params &b)
{
    return a.alpha0 == b.alpha0 && a.f_alpha == b.f_alpha && a.dt_max == b.dt_max &&
           a.f_inc == b.f_inc && a.f_dec == b.f_dec && a.nmin == b.nmin && a.mass == b.mass &&
           a.fcut == b.fcut && a.max_steps == b.max_steps && a.criterion == b.criterion;
}

// The FIRE scalars of a fresh relax: slot 0 = (dt_max, alpha0, 0, not done,
// step 0), slot 1 and the result cleared.
hipError_t reset_state(nr_composite &c, const nr_relax_params &p)
{
    double h[kStWords] = {0};
    h[kStDt] = p.dt_max;
    h[kStAlpha] = p.alpha0;
    hipError_t e = hipMemcpyAsync(c.state.p, h, sizeof(h), hipMemcpyHostToDevice, c.st);
    if (e == hipSuccess) e = hipStreamSynchronize(c.st);
    return e;
}

hipError_t capture_chunk(nr_composite &c)
{
    // nothing may still run the graph being replaced (a relax that returned
    // early can leave replays queued on the stream)
    hipError_t e = hipStreamSynchronize(c.st);
    if (e != hipSuccess) return e;
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c.exec = nullptr;
    c.graph = nullptr;
    c.have_graph = false;
    e = hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    // the chunk starts reading slot 1 (the first drift writes it)
    for (int k = 0; k < kMeshChunk && e == hipSuccess; k++)
        e = (hipError_t)launch_mesh_step(c.md, (k + 1) & 1, 0, c.st);
    hipError_t e2 = hipStreamEndCapture(c.st, &c.graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0);
    if (e == hipSuccess) c.have_graph = true;
    return e;
}

// A point array of the device (2 doubles per point, device order) into
// `out` in the file's point order.
int to_file_order(nr_composite &c, const double *dev, double *out)
{
    const size_t n = (size_t)c.md.n;
    std::vector<double> d(2 * n);
    HIP_TRY(hipMemcpyAsync(d.data(), dev, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c.st));
    HIP_TRY(hipStreamSynchronize(c.st));
    for (size_t i = 0; i < n; i++) {
        out[2 * i] = d[2 * (size_t)c.dev_of[i]];
        out[2 * i + 1] = d[2 * (size_t)c.dev_of[i] + 1];
    }
    return NR_OK;
}

int energies(nr_composite &c, double out[3])
{
    std::vector<double> h(3 * (size_t)c.energy_blocks);
    HIP_TRY((hipError_t)launch_mesh_energy(c.md, c.epart.p, c.energy_blocks, c.st));
    HIP_TRY(hipMemcpyAsync(h.data(), c.epart.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                           c.st));
    HIP_TRY(hipStreamSynchronize(c.st));
    out[0] = out[1] = out[2] = 0.0;
    for (int b = 0; b < c.energy_blocks; b++)
        for (int k = 0; k < 3; k++) out[k] += h[3 * (size_t)b + k];
    return NR_OK;
}

int relax(nr_composite &c, const nr_relax_params &p, nr_report &rep, double *e_gel)
{
    if (p.method != NR_METHOD_FIRE)
        return fail(NR_E_ARG, "composite relax supports FIRE only (the reference's method)");
    if (!(p.mass > 0) || !(p.dt_max > 0) || p.nmin < 0 || p.max_steps < 0)
        return fail(NR_E_ARG, "invalid FIRE parameters");
    HIP_TRY(hipSetDevice(c.device));
    set_fire(c.md, p);
    if (c.graphs && (!c.have_graph || !same_params(c.captured, p))) {
        HIP_TRY(capture_chunk(c));
        c.captured = p;
    }
    HIP_TRY(reset_state(c, p));
    HIP_TRY(hipMemsetAsync(c.v.p, 0, sizeof(double) * 2 * (size_t)c.md.n, c.st));
    HIP_TRY(hipEventRecord(c.e0, c.st));
    HIP_TRY((hipError_t)launch_mesh_step(c.md, 0, 1, c.st));
    double h[kStWords];
    const long max_replays = (long)p.max_steps / kMeshChunk + 3;
    long replays = 0;
    for (;;) {
        HIP_TRY(hipMemcpyAsync(h, c.state.p, sizeof(h), hipMemcpyDeviceToHost, c.st));
        HIP_TRY(hipStreamSynchronize(c.st));
        if (h[kStSlot + kStDone] != 0.0) break;
        if (replays++ > max_replays)
            return fail(NR_E_TIMEOUT, "composite relax did not finish within its step bound");
        if (c.graphs) {
            HIP_TRY(hipGraphLaunch(c.exec, c.st));
        } else {
            for (int k = 0; k < kMeshChunk; k++)
                HIP_TRY((hipError_t)launch_mesh_step(c.md, (k + 1) & 1, 0, c.st));
        }
    }
    HIP_TRY(hipEventRecord(c.e1, c.st));
    HIP_TRY(hipEventSynchronize(c.e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c.e0, c.e1));
    c.last_ms = ms;
    const double *r = h + 2 * kStSlot;
    double e[3];
    int rc = energies(c, e);
    if (rc) return rc;
    std::memset(&rep, 0, sizeof(rep));
    rep.rms_force = r[kResRms];
    rep.max_force = r[kResFmax];
    rep.e_stretch = e[0];
    rep.e_bend = e[1];
    rep.steps = (int)r[kResSteps];
    rep.status = (int)r[kResStatus];
    rep.fp64_steps = rep.steps;
    if (e_gel) *e_gel = e[2];
    return NR_OK;
}

}  // namespace
}  // namespace nr

// ---- C ABI ----------------------------------------------------------------

extern "C" int nr_mesh_read(const char *path, nr_mesh **out)
{
    return nr::guarded(NR_E_OPEN, [&]() -> int {
        if (!path || !out) return nr::fail(NR_E_ARG, "null argument");
        *out = nullptr;
        auto m = std::make_unique<nr_mesh>();
        std::string err;
        int rc = nr::read_mesh(path, m->m, err);
        if (rc) return nr::fail(rc, err);
        *out = m.release();
        return NR_OK;
    });
}

extern "C" int nr_mesh_lattice(int L, double p, uint64_t seed, nr_mesh **out)
{
    if (!out) return nr::fail(NR_E_ARG, "null argument");
    *out = nullptr;
    auto m = std::make_unique<nr_mesh>();
    std::string err;
    int rc = nr::make_mesh_lattice(L, p, seed, m->m, err);
    if (rc) return nr::fail(rc, err);
    *out = m.release();
    return NR_OK;
}

extern "C" int nr_mesh_write_nmc(const nr_mesh *m, const char *path)
{
    if (!m || !path) return nr::fail(NR_E_ARG, "null argument");
    int rc = nr::write_nmc(m->m, path);
    return rc ? nr::fail(rc, std::string("cannot write ") + path) : NR_OK;
}

extern "C" int nr_mesh_get_info(const nr_mesh *m, nr_mesh_info *info)
{
    if (!m || !info) return nr::fail(NR_E_ARG, "null argument");
    info->n_points = m->m.net.n;
    info->n_bonds = m->m.net.e();
    info->n_facets = m->m.nf();
    info->n_twins = m->m.np();
    info->offset = m->m.offset;
    info->minx = m->m.minx;
    info->maxx = m->m.maxx;
    info->miny = m->m.net.miny;
    info->maxy = m->m.net.maxy;
    return NR_OK;
}

extern "C" int nr_mesh_get_arrays(const nr_mesh *m, double *pos, int *facets, double *areas,
                                  int *twins)
{
    if (!m) return nr::fail(NR_E_ARG, "null argument");
    const nr::Mesh &s = m->m;
    if (pos) std::memcpy(pos, s.net.pos.data(), sizeof(double) * s.net.pos.size());
    if (facets) std::memcpy(facets, s.facets.data(), sizeof(int) * s.facets.size());
    if (areas) std::memcpy(areas, s.areas.data(), sizeof(double) * s.areas.size());
    if (twins) std::memcpy(twins, s.twins.data(), sizeof(int) * s.twins.size());
    return NR_OK;
}

extern "C" void nr_mesh_free(nr_mesh *m) { delete m; }

extern "C" int nr_composite_create(const nr_mesh *m, const nr_gel_params *g, int device,
                                   nr_composite **out)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (!m || !g || !out) return nr::fail(NR_E_ARG, "null argument");
        *out = nullptr;
        auto c = std::make_unique<nr_composite>();
        c->mesh = m->m;
        c->gel = *g;
        c->device = device;
        const char *g_env = std::getenv("NRELAX_GRAPHS");
        c->graphs = !(g_env && g_env[0] == '0');
        std::string err;
        int rc = nr::mesh_setup(c->mesh, g->ks, g->kb, g->bmod, g->dilation, c->setup, err);
        if (rc) return nr::fail(rc, err);
        if (!(g->smod >= 0)) return nr::fail(NR_E_ARG, "smod must be >= 0");
        HIP_TRY(hipSetDevice(device));
        HIP_TRY(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&c->e0));
        HIP_TRY(hipEventCreate(&c->e1));
        rc = nr::composite_upload(*c);
        if (rc) return rc;
        *out = c.release();
        return NR_OK;
    });
}

extern "C" int nr_composite_get_info(const nr_composite *c, nr_composite_info *info)
{
    if (!c || !info) return nr::fail(NR_E_ARG, "null argument");
    info->width = c->setup.width;
    info->range = c->setup.range;
    info->pref_d = c->setup.pref_d;
    info->net_bmod = c->setup.net_bmod;
    info->n_triples = c->md.nt;
    info->n_bonds = c->md.nb;
    info->n_facets = c->md.nf;
    info->n_points = c->md.n;
    return NR_OK;
}

extern "C" int nr_composite_affine(nr_composite *c, double yrange, double xdisp, double ydisp)
{
    if (!c) return nr::fail(NR_E_ARG, "null argument");
    if (!(yrange != 0.0) || !std::isfinite(xdisp) || !std::isfinite(ydisp))
        return nr::fail(NR_E_ARG, "invalid affine step");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY((hipError_t)nr::launch_mesh_affine(c->x.p, c->md.n, c->mesh.net.miny, yrange, xdisp,
                                               ydisp, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    return NR_OK;
}

extern "C" int nr_composite_relax(nr_composite *c, const nr_relax_params *p, nr_report *rep,
                                  double *e_gel)
{
    if (!c || !p || !rep) return nr::fail(NR_E_ARG, "null argument");
    return nr::relax(*c, *p, *rep, e_gel);
}

extern "C" int nr_composite_get_positions(nr_composite *c, double *pos)
{
    if (!c || !pos) return nr::fail(NR_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    return nr::to_file_order(*c, c->x.p, pos);
}

extern "C" int nr_composite_set_positions(nr_composite *c, const double *pos)
{
    if (!c || !pos) return nr::fail(NR_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    const size_t n = (size_t)c->md.n;
    std::vector<double> d(2 * n);
    for (size_t i = 0; i < n; i++) {
        d[2 * (size_t)c->dev_of[i]] = pos[2 * i];
        d[2 * (size_t)c->dev_of[i] + 1] = pos[2 * i + 1];
    }
    HIP_TRY(hipMemcpyAsync(c->x.p, d.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    return NR_OK;
}

// Force (clamped, twins summed) and {E_stretch, E_bend, E_gel} at the
// current positions, by the relax's own kernels.
extern "C" int nr_composite_eval(nr_composite *c, double *force, double *energies)
{
    if (!c) return nr::fail(NR_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    nr_relax_params p{};
    p.dt_max = 0.1;
    p.alpha0 = 0.1;
    p.mass = 1.0;
    p.max_steps = 0;  // the first-drift launch then stops at once: x stays
    nr::set_fire(c->md, p);
    HIP_TRY(nr::reset_state(*c, p));
    HIP_TRY(hipMemsetAsync(c->v.p, 0, sizeof(double) * 2 * (size_t)c->md.n, c->st));
    HIP_TRY((hipError_t)nr::launch_mesh_step(c->md, 0, 1, c->st));
    if (force) {
        const int rc = nr::to_file_order(*c, c->f.p, force);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(c->st));
    if (energies) {
        int rc = nr::energies(*c, energies);
        if (rc) return rc;
    }
    return NR_OK;
}

extern "C" double nr_composite_last_ms(const nr_composite *c) { return c ? c->last_ms : 0.0; }

extern "C" void nr_composite_destroy(nr_composite *c)
{
    if (c) (void)hipSetDevice(c->device);
    delete c;
}

// ---- the composite program's batch protocol ------------------------------

namespace nr {
namespace {

struct MeshBatchSpec {
    BatchSpec b;  // lines 1 and 3-6 as the network program's
    nr_gel_params gel{};
};

// parse_batch_file :413-500: as the network program's, except line 2 is
// "ks kb bmod [smod [dilation]]" (at least three numbers, none negative).
int parse_mesh_batch(const std::string &path, MeshBatchSpec &s, std::string &err)
{
    int rc = parse_batch_file(path, s.b, err);
    if (rc) return rc;
    std::ifstream in(path);
    std::string line;
    int got = 0;
    while (got < 2 && std::getline(in, line))
        if (!line.empty() && line[0] != '#') got++;
    nr_gel_params &g = s.gel;
    g = nr_gel_params{};
    int nr = std::sscanf(line.c_str(), "%lf %lf %lf %lf %lf", &g.ks, &g.kb, &g.bmod, &g.smod,
                         &g.dilation);
    if (nr < 3 || g.ks < 0 || g.kb < 0 || g.bmod < 0 || g.smod < 0 || g.dilation < 0) {
        err = "Error: invalid format for mechanical attributes.";
        return NR_E_BATCH;
    }
    return NR_OK;
}

void log_mesh_result(FILE *lf, const nr_report &r, double fcut)
{
    std::printf("Steps: %d, RMS Force: %12.8le\n", r.steps, r.rms_force);
    if (!lf) return;
    std::fprintf(lf, r.rms_force <= fcut ? "Convergence reached: " : "Convergence not reached: ");
    std::fprintf(lf, "Steps: %d, RMS Force: %12.8le\n", r.steps, r.rms_force);
}

// do_relaxation_run :1095-1391 on one composite.  Compression: the first
// step applies the minimum, each later one the increment, the height shrinking
// by each (:1283-1327); shear likewise from the minimum (:1335-1377).
int run_mesh_protocol(const JobFiles &jf, const MeshBatchSpec &s, bool strain_mode,
                      int device, std::string &err)
{
    const BatchSpec &b = s.b;
    if ((b.cmax > 0 && !(b.cinc > 0)) || (b.smax > 0 && !(b.sinc > 0))) {
        err = "a ramp with a positive maximum needs a positive increment";
        return NR_E_BATCH;
    }
    clock_t ct = std::clock();
    nr_mesh *mp = nullptr;
    int rc = nr_mesh_read(jf.dat.c_str(), &mp);
    if (rc) {
        err = nr_last_error();
        return rc;
    }
    std::unique_ptr<nr_mesh, void (*)(nr_mesh *)> mesh(mp, nr_mesh_free);
    nr_composite *cp = nullptr;
    rc = nr_composite_create(mesh.get(), &s.gel, device, &cp);
    if (rc) {
        err = nr_last_error();
        return rc;
    }
    std::unique_ptr<nr_composite, void (*)(nr_composite *)> comp(cp, nr_composite_destroy);
    const int n = mesh->m.net.n;
    double range = comp->setup.range;
    const double smult = strain_mode ? range : 1.0;
    const int digits = dec_digits(1 + (b.sinc > 0 ? (int)((b.smax - b.smin) / b.sinc) : 0));
    nr_relax_params p{};
    p.alpha0 = b.fp[0];
    p.f_alpha = b.fp[1];
    p.dt_max = b.fp[2];
    p.f_inc = b.fp[3];
    p.f_dec = b.fp[4];
    p.nmin = b.nmin;
    p.mass = b.mass;
    p.fcut = b.fcut;
    p.max_steps = b.nmax;
    p.criterion = NR_CRIT_RMS;
    p.method = NR_METHOD_FIRE;
    std::cout << "Preferred dilation: " << comp->setup.pref_d << "\n";
    std::cout.flush();
    FILE *lf = std::fopen(jf.log.c_str(), "w");
    FILE *ef = std::fopen(jf.edata.c_str(), "w");
    std::vector<double> pos(2 * (size_t)n);
    nr_report r{};
    double eg = 0.0, cur = 0.0;
    double mn = b.cmin * smult, mx = b.cmax * smult, inc = b.cinc * smult;
    if (mx > 0) {
        rc = nr_composite_affine(cp, range, 0.0, -mn);
        range -= mn;
        for (cur = mn; rc == NR_OK && cur <= mx + 1e-10; cur += inc) {
            rc = nr_composite_relax(cp, &p, &r, &eg);
            if (rc) break;
            log_mesh_result(lf, r, p.fcut);
            if (ef)
                std::fprintf(ef, "%12.0lf\t%12.10lf\t%12.8le\t%12.8le\t%12.8le\t%12.8le\n", 0., cur,
                             r.e_stretch, r.e_bend, eg, r.e_stretch + r.e_bend + eg);
            if (cur < mx) {
                rc = nr_composite_affine(cp, range, 0.0, -inc);
                range -= inc;
            }
        }
    }
    const double max_comp = mx > mn ? cur - inc : mx;
    mn = b.smin * smult;
    mx = b.smax * smult;
    inc = b.sinc * smult;
    int count = 1;
    if (rc == NR_OK && mx > 0) {
        rc = nr_composite_affine(cp, range, mn, 0.0);
        for (cur = mn; rc == NR_OK && cur <= mx + 1e-10; cur += inc) {
            rc = nr_composite_relax(cp, &p, &r, &eg);
            if (rc) break;
            log_mesh_result(lf, r, p.fcut);
            if (ef)
                std::fprintf(ef, "%12.8lf\t%12.8lf\t%12.8le\t%12.8le\t%12.8le\t%12.8le\n", cur,
                             max_comp, r.e_stretch, r.e_bend, eg, r.e_stretch + r.e_bend + eg);
            rc = nr_composite_get_positions(cp, pos.data());
            if (rc) break;
            FILE *pf = std::fopen(report_name(jf.disp, digits, count++, "pos").c_str(), "w");
            if (pf) {
                for (int i = 0; i < n; i++)
                    std::fprintf(pf, "%2.10le\t%2.10le\n", pos[2 * i], pos[2 * i + 1]);
                std::fclose(pf);
            }
            if (cur < mx) rc = nr_composite_affine(cp, range, inc, 0.0);
        }
    }
    if (rc) err = nr_last_error();
    ct = std::clock() - ct;
    if (lf) {
        std::fprintf(lf, "Elapsed time: %f seconds\n", ((float)ct) / CLOCKS_PER_SEC);
        std::fclose(lf);
    }
    if (ef) std::fclose(ef);
    return rc;
}

}  // namespace
}  // namespace nr

// relaxation_batch :1393-1421: every <base>*.nmc of the batch's directory.
extern "C" int nr_mesh_relaxation_batch(const char *batch_file, int strain_mode, int device)
{
    return nr::guarded(NR_E_BATCH, [&]() -> int {
        std::string err;
        nr::MeshBatchSpec s;
        if (!batch_file) return NR_E_ARG;
        int rc = nr::parse_mesh_batch(batch_file, s, err);
        if (rc) {
            std::fprintf(stderr, "%s\nParsing of the batch file failed.\n", err.c_str());
            return rc;
        }
        std::vector<nr::JobFiles> jobs;
        rc = nr::file_lists(s.b.base, s.b.tag, jobs, err, "nmc");
        if (rc) {
            std::fprintf(stderr, "%s\nFile lists could not be read.\n", err.c_str());
            return rc;
        }
        int first_err = NR_OK;
        for (const nr::JobFiles &j : jobs) {
            rc = nr::run_mesh_protocol(j, s, strain_mode != 0, device, err);
            if (rc) {
                std::fprintf(stderr, "%s: %s\n", j.dat.c_str(), err.c_str());
                if (!first_err) first_err = rc;
            }
        }
        return first_err;
    });
}
