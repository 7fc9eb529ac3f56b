// Structures shared by the host driver and the gfx950 kernels.
#pragma once

#include <cstdint>

#include "plan.hpp"
#include "../../include/nrelax.h"

namespace nr {

// Per-network parameters and persistent state on the device.
struct NetDev {
    int node_base, n_nodes, tile_base, n_tiles;
    int parity;       // which state buffer holds the current x, v, F
    int crit, nmin, max_steps;
    int method, pad0, pad1, pad2;
    double width, height, le_shift, miny, yrange;
    double ks, kb;
    double a0, fa, dtmax, finc, fdec, mass, fcut;
    double pend_x, pend_y;  // affine step applied by the next launch
};

struct NetOut {
    double rms, fmax, es, eb;
    int steps, status, fp64_steps, pad;
};

enum KernelMode { kModeRelax = 0, kModeEval = 1 };

struct KArgs {
    double2 *x[2], *v[2], *f[2];
    const unsigned char *mask;
    const int *brow, *trow;
    const uint32_t *bcode;
    const double *blen;
    const uint64_t *tcode;
    const double *tkb;
    const int *halo;
    const TileDev *tiles;
    NetDev *nets;
    NetOut *out;
    const int *jobs;
    int n_jobs;
    int team;          // workgroups per network
    int xcd_local;     // teams formed from workgroups of one XCD
    int has_tkb;
    int max_local, max_nb, max_nt;
    int mode;
    int *queue;        // job counter
    double *partials;  // [n_teams][team][4]
    unsigned *bar;     // per team: [team*64] count, [team*64+32] generation
    int *mailbox;      // per team, stride 32
    unsigned *err;     // nonzero after a bounded wait expired
    long long timeout_ticks;
};

// Host entry: launches the persistent relax kernel.  grid workgroups of
// `block` threads, nodes-per-thread npt; returns a hipError_t value.
int launch_relax(const KArgs &a, int npt, int grid, int block, size_t lds,
                 bool cooperative, void *stream);
int max_resident_blocks(int npt, int block, size_t lds, int *per_cu, int *n_cu);

}  // namespace nr
