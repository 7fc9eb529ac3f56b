// Structures shared by the host driver and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/nrelax.h"
#include "plan.hpp"

namespace nr {

// Per-network parameters and persistent state on the device.
struct NetDev {
    int node_base, n_nodes, tile_base, n_tiles;
    int parity;       // which state buffer holds the current x, v, F
    int crit, nmin, max_steps;
    int method;
    int inc_count, inc_base;  // strain increments of this network in KArgs::incs
    int out_base;             // its per-increment reports in KArgs::inc_out
    double width, height, le_shift, miny, yrange;
    double ks, kb;
    double a0, fa, dtmax, finc, fdec, mass, fcut;
    double pend_x, pend_y;  // affine step applied by the next launch
};

// Resumable relax state of one network (all zero = start a fresh relax).
// A launch with a step budget stops a relax mid-way and saves here what the
// next launch needs to continue it with the same arithmetic.
enum RunPhase { kRunFresh = 0, kRun32 = 1, kRun64 = 2, kRunCg = 3 };
struct NetRun {
    int phase;
    int step, steps32;       // FIRE steps of the current increment, FP32 stage
    int since;               // FIRE steps since the power was last negative
    int since_best;          // FP32 stage: steps without a new minimum
    int mix, s;              // deferred mix of the next pass; FP32 parity (CG:
                             // parity of the direction buffer)
    int inc_next;            // increments of the list already started
    double alpha, dt;        // FIRE mixing factor and time step (CG: next beta,
                             // line-search step length)
    double alpha_old, cmix;  // deferred mix coefficients (CG: |d|^2, F.d of
                             // the last accepted trial)
    double best, fswitch;    // FP32 stage: least max |F|^2, hand-over threshold
    double ff, fmax2;        // |F|^2 and max node |F|^2 after the last pass
    long long t_start;       // device clock when the increment began
    int slice;               // steps this network may take in the current launch
                             // (set by the host per launch; 0 = KArgs::slice_steps)
    int pad;
};

struct NetOut {
    double rms, fmax, es, eb;
    int steps, status, fp64_steps, pad;
    long long t_start, t_done;  // 100 MHz device clock at job start / finish
    double sxy;                 // shear stress: shear virial / box area
};

enum KernelMode { kModeRelax = 0, kModeEval = 1 };

// Role of one workgroup of the persistent grid (host-built table).
struct BlockRole {
    int team;     // team id (-1: idle)
    int member;   // index within the team
    int g;        // team size
    int cls;      // job class (team-size class) the team serves
    int part;     // first partial-sum slot of the team (one per member)
    int tree;     // first barrier/mailbox id of the team's split tree (2g - 1 ids:
                  // the team, its halves, their halves, ... in heap order)
    int pad[2];
};

constexpr int kMaxLog2Team = 16;

struct KArgs {
    double2 *x[2], *v[2], *f[2];
    // mixed precision: FP32 displacement from the FP64 anchor, velocity, force
    float2 *u32[2], *v32[2], *f32[2];
    int mixed;               // relax with an FP32 stage and an FP64 polish
    double switch_ratio;     // FP32 stage ends at |F|max < switch_ratio * initial |F|max
    int stall_steps;         // ... or after this many steps without a new minimum
    const unsigned char *mask;
    const uint4 *blob;
    const TileDev *tiles;
    NetDev *nets;
    NetOut *out;
    const BlockRole *roles;   // per workgroup
    const int *class_jobs;    // networks of each class, largest first
    const int *class_off;     // [n_classes + 1]
    int *queues;              // per class job counter, stride 32
    int tile_cap;            // LDS bytes of each workgroup's tile area (blobs + work)
    int pad_k[2];
    int mode;
    unsigned long long *partials;  // per team member kPartWords tagged words (zeroed per launch)
    unsigned long long *bar;  // per team kBarStride words: see team_arrive
    int *mailbox;             // per team, stride 32
    unsigned *err;            // nonzero after a bounded wait expired
    long long timeout_ticks;
    int cls_of_log2g[kMaxLog2Team];  // class serving teams of 2^i workgroups, -1 none
    NetRun *run;              // per network resumable state
    const double2 *incs;      // strain increments (xdisp, ydisp)
    NetOut *inc_out;          // one report per increment, at NetDev::out_base
    int slice_steps;          // FIRE steps per network per launch, 0 = no limit
};

// LDS cache, per workgroup, of the tile descriptors of the first kTileCache
// tiles a team member relaxes in its current network (each with where its
// blob is held when the blobs of all of them stay resident in LDS, or else
// where its halo ids are) and of their halo node ids (a pool of kCachePool
// ids filled in tile order; a tile whose list no longer fits is not
// cached), placed before the tile area: with the halo ids at hand a tile
// pass issues its blob, own-node and halo-node loads at once.  Sized for
// 8-16 tiles of 256 nodes (~125 halo ids each) per member.
constexpr int kTileCache = 16;
constexpr int kCachePool = 1808;
constexpr int kCacheBytes = kTileCache * (int)sizeof(TileDev) + kCachePool * 4;
static_assert(kCacheBytes % 16 == 0, "tile area alignment");
// LDS of a workgroup outside the tile area: the cache and static scratch.
constexpr int kLdsReserve = kCacheBytes + 1024;
// LDS of one workgroup: two 256-thread workgroups share a CU's 160 KB, a
// 512-thread workgroup has it alone.
inline int lds_per_workgroup(int block) { return block == 256 ? 80 * 1024 : 156 * 1024; }

// Partial-sum record of a team member: two step parities x 8 words, each
// word = step tag (high 32 bits) | one half of a double (low 32 bits).
constexpr int kPartWords = 16;
// Every member reads every member's record, four lanes per record, at most
// poll_sweeps(npt) records per lane (fewer in the kernels with two nodes per
// thread, which have fewer registers to spare): a team has at most
// poll_sweeps(npt) * block / 4 members.
constexpr int poll_sweeps(int npt) { return npt == 1 ? 8 : 4; }
inline int max_team(int block, int npt) { return poll_sweeps(npt) * block / 4; }

// Team barrier words per team: [0] root arrivals, [16] generation,
// [32 + 16 g] arrivals of group g (groups of kBarGroup members).
constexpr int kBarGroup = 8;
constexpr int kBarStride = 32 + 16 * 64;

// Host entry: launches the persistent relax kernel (grid workgroups of
// `block` threads, npt own nodes per thread).  Returns a hipError_t value.
// method: NR_METHOD_FIRE or NR_METHOD_CG (every network of the launch).
int launch_relax(const KArgs &a, int block, int npt, int method, int grid, size_t lds,
                 bool cooperative, void *stream);
int max_resident_blocks(int block, int npt, size_t lds, int *per_cu, int *n_cu);
// pend of network k += d[k] on the device (stream-ordered, graph-capturable).
int launch_strain_step(NetDev *nets, int count, const double2 *d, void *stream);
// Zeroes `bytes` (a multiple of 8) with a kernel (graph-capturable).
int launch_zero(void *p, size_t bytes, void *stream);

// Post-relax reductions of one network (analysis.hip): per-workgroup partial
// sums over kAnalysisBlocks workgroups into part[0 .. kAnalysisBlocks)
// (non-affinity) and part[kAnalysisBlocks + 3 w + {0, 1, 2}] (axial energy,
// strain sum, max |strain|); strain[e] per bond.  bonds: (i, j, wx, wy) in
// the network's device node indices, edge order.
constexpr int kAnalysisBlocks = 256;

// Small-strain linear solve (linear_kernels.hip): a batch of independent
// systems (segments).  Segment i solves over rows [row0, row0 + n) of the
// batch's vectors with the CSR matrix whose n + 1 row offsets start at
// rowptr[rp0] (offsets relative to nz0 in col and val, columns local to the
// segment) and whose free-DOF mask starts at free_dof[fr0] (rows of clamped
// DOFs masked), served by workgroups [blk0, blk0 + nblk) of each launch
// (nblk = lin_segment_blocks(n)); seg_of maps a workgroup to its segment.
// Segments of one network and stiffness share its matrix.  Iteration k of a
// chunk (k from 0, chunks of an even length) of CG on K_rr w = b, the
// direction alternating between p[k & 1] (read) and p[(k & 1) ^ 1]
// (written), kLinScalars scalars per segment (see linear_kernels.hip),
// partials of p.q then r.r in part[0 .. 2 blocks); a chunk ends with
// launch_cg_chunk_end.
struct LinSeg {
    int row0, n, blk0, nblk;
    int rp0, fr0;
    int max_iter;  // iterations the segment may take
    int pad;
    long long nz0;
};
constexpr int kLinBlocks = 512;  // most workgroups of one segment
constexpr int kLinScalars = 9;
int lin_segment_blocks(int n);
int launch_cg_iteration(const LinSeg *segs, const int *seg_of, int blocks, const int *rowptr,
                        const int *col, const double *val, const unsigned char *free_dof,
                        double *w, double *r, double *const p[2], double *q, double *part,
                        double *sc, int k, void *stream);
int launch_cg_chunk_end(double *sc, int count, int chunk, void *stream);
// strain may be null (the bond strains are then not stored).
int launch_analysis(const double2 *x, const double2 *x0, int n, double miny, double shear,
                    double axial, const int4 *bonds, int ne, double W, double H, double shift,
                    double *strain, double *part, void *stream);
// Fixed workspace of the allocation-free analysis (analysis.hip): partials,
// then kNoAllocMaxBonds bond strains, then a staged list of as many bonds.
constexpr int kNoAllocMaxBonds = NR_NOALLOC_MAX_BONDS;
constexpr size_t kNoAllocScratchDoubles = 4 * kAnalysisBlocks + 3 * (size_t)kNoAllocMaxBonds;
// Device address of that workspace on `device`, which must be current.
int analysis_scratch(int device, double **ptr);

}  // namespace nr
