// Device layout and launches of the network + FEM-gel composite relax
// (mesh_kernels.hip, composite.cpp).
#pragma once

#include <hip/hip_runtime.h>

namespace nr {

// Scalar state: two parity slots of kStSlot doubles, then the result.
constexpr int kStSlot = 8;
constexpr int kStDt = 0, kStAlpha = 1, kStSince = 2, kStDone = 3, kStStep = 4;
constexpr int kResRms = 0, kResFmax = 1, kResSteps = 2, kResStatus = 3;
constexpr int kStWords = 3 * kStSlot;
constexpr int kMeshRunning = -1, kMeshConverged = 0, kMeshNotConverged = 1, kMeshNonfinite = 2;
// Most workgroups of a point or element launch (grid-stride beyond).
constexpr int kMeshMaxBlocks = 2048;

// A tile of the fused element + gather kernel (mesh_tile): a run of at most
// kMeshTilePoints points in Morton order (its own points), every element
// incident to an own point or to an own point's twin, and the points those
// elements reach (own first, then halo).  Its blob (16-byte aligned, copied
// into LDS) holds, each section 16-byte aligned:
//   ids   int32[n_own + n_halo]  point ids, own then halo
//   ioff  u16[n_own + 1]         entry list start of each own point
//   imid  u16[n_own]             start of the twin's entries in that list
//   mask  u8[n_own]              clamp code | twin's clamp code << 2 | has twin << 4
//   inc   u16[n_inc]             (force vector << 1) | negate: the point's own
//                                bond, triple and facet entries in element
//                                order, then its twin's
// Force vectors in LDS: bonds [0, nb) (force on the first point), triple
// ends a [nb, nb + nt) and b [nb + nt, nb + 2 nt), facet corner k
// [nb + 2 nt + k nf, ...): the tile's elements in global element order.
// Element records are read from global memory (tile-major arrays).
constexpr int kMeshTilePoints = 256;  // at most one own point per thread
// Element records a thread of mesh_tile holds in registers (256 threads):
// a tile has at most 256 times as many of each type.
constexpr int kTileBondsPerThread = 3, kTileTriplesPerThread = 2, kTileFacetsPerThread = 3;
// ... and points (own + halo) of the tile whose ids and positions it loads,
// and 16-byte rows of the tile's blob it stages
constexpr int kTileLocalPerThread = 2, kTileBlobRows = 4;
struct MeshTile {
    int blob16, blob_bytes, n_own, n_halo;
    int n_inc, nb, nt, nf;
    int b0, t0, f0, pad;  // first record of each element type
};
static_assert(sizeof(MeshTile) == 48, "MeshTile layout");

struct MeshBlob {
    int ids, ioff, imid, mask, inc, bytes;
};

__host__ __device__ inline int mesh_align16(int x) { return (x + 15) & ~15; }

__host__ __device__ inline MeshBlob mesh_blob(int n_own, int n_halo, int n_inc)
{
    MeshBlob b;
    int o = 0;
    b.ids = o;  o += mesh_align16(4 * (n_own + n_halo));
    b.ioff = o; o += mesh_align16(2 * (n_own + 1));
    b.imid = o; o += mesh_align16(2 * n_own);
    b.mask = o; o += mesh_align16(n_own);
    b.inc = o;  o += mesh_align16(2 * n_inc);
    b.bytes = o;
    return b;
}

struct MeshDev {
    int n;                       // points
    int nb, nt, nf;              // bonds, triples, facets
    double *x, *v, *f;           // 2n each
    // Each element writes its forces into its own slots, in element order
    // (contiguous writes); a point sums the slots that name it.
    const int2 *bond;            // (i, j)
    const double2 *bond_geo;     // (rest length, x image offset of j)
    const int4 *tri;             // (hinge, end a, end b, 0)
    const double2 *tri_off;      // x image offsets of the two arms
    const int4 *facet;           // (p1, p2, p3, 0)
    const double *inv;           // 6 per facet
    const double *area;          // per facet
    double2 *slot;               // element forces: bonds [0, nb), triple ends a then
                                 // b (nt each), facet corners 0, 1, 2 (nf each)
    // Slot entries of each point, (slot << 1) | negate, in element order, in
    // three sections (bonds, triples, facets), each stored column-major:
    // entry k of point i in section s at ent[sec[s] + k * n + i], so the
    // k-th entries of consecutive points are consecutive words.
    const int *ent;
    const int *cnt;              // n: entries per section, 8 bits each
    long sec[3];                 // section offsets in ent
    const int *twin;             // n: twin point or -1; null when twins are not coupled
    const unsigned char *fcode;  // n clamp codes
    double *part;                // 4 per gather workgroup
    double *tot;                 // 4: the partials' sums (v.F, v.v, F.F, max |F|^2)
    unsigned *ticket;            // gather workgroups done (0 between launches)
    double *st;                  // kStWords
    int gather_blocks;
    // fused element + gather kernel (mesh_tile)
    int n_tiles;
    int tile_grid;               // workgroups of mesh_tile (at most the resident ones)
    int tile_lds;                // dynamic LDS of mesh_tile
    const MeshTile *tiles;
    const uint4 *tblob;
    const unsigned *tb_ij;       // bond: first | second << 16 (tile-local)
    const double2 *tb_geo;       // bond: (rest length, x image offset of the second)
    const uint2 *tt_idx;         // triple: hinge | end a << 16, end b
    const double2 *tt_off;       // triple: x image offsets of the two arms
    const uint2 *tf_idx;         // facet: corner 0 | corner 1 << 16, corner 2
    const double2 *tf_m;         // facet: inverse rest matrix, 3 per facet
    const double *tf_area;       // facet: area
    double ks, kb, smod, bmod, pref_d;
    double alpha0, f_alpha, dt_max, f_inc, f_dec, mass, fcut, sqrt_dof;
    int nmin, max_steps, criterion, pad;
};

int mesh_element_blocks(const MeshDev &m);
int mesh_point_blocks(int n);
// phase 0: one FIRE step reading state slot par; 1: the initial force and
// the first drift (reads slot par, writes par ^ 1).
int launch_mesh_step(const MeshDev &m, int par, int phase, void *stream);
int launch_mesh_affine(double *x, int n, double miny, double yrange, double xdisp, double ydisp,
                       void *stream);
int launch_mesh_energy(const MeshDev &m, double *part, int blocks, void *stream);

}  // namespace nr
